// Input formats of minipoa-rs's `io` module (src/io.rs is not in the
// reference's tree; its behaviour is known from tests/io_cli.rs:34-87,
// 374-419, 563-610 and README.md): FASTA/FASTQ records, gzip input, and the
// GFA graph the -i option continues (src/lib.rs:214-216).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "graph.hpp"

namespace mpoa {

// src/lib.rs Sequence: name, comment, bases and (FASTQ) qualities.
struct SeqRecord {
    std::string name, comment;
    std::vector<uint8_t> seq, qual;
};

// Decompresses data when it starts with the gzip magic (every concatenated
// member), else copies it. Empty string on success, else the error.
std::string gunzip_if_gzip(const uint8_t *data, size_t len, std::vector<uint8_t> &out);

// io::parse_sequences: FASTA or FASTQ text (tests/io_cli.rs:34-75).
std::string parse_sequences(const uint8_t *text, size_t len, std::vector<SeqRecord> &out);

// io::read_gfa: a forward, acyclic GFA1 graph (H/S/L/P records, explicit
// segment sequences, zero overlaps) as a POA graph plus one record per path.
// Segments become one node per base in file order, chained; links and paths
// add edges; paths are replayed as reads (edge weights count them, links no
// path uses weigh 1); nodes without a predecessor hang off the source and
// nodes without a successor lead to the sink. Every node is its own
// alignment column (GFA does not carry column groups). The graph is
// toposorted.
std::string read_gfa(const Params &p, const uint8_t *text, size_t len, Graph &g,
                     std::vector<SeqRecord> &records);

}  // namespace mpoa
