// Sequence and graph input (the reference's io module, known from
// tests/io_cli.rs and README.md; src/io.rs is not in its tree).
#include "io.hpp"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cstring>
#include <unordered_map>

namespace mpoa {

std::string gunzip_if_gzip(const uint8_t *data, size_t len, std::vector<uint8_t> &out) {
    out.clear();
    if (len < 2 || data[0] != 0x1f || data[1] != 0x8b) {
        out.assign(data, data + len);
        return {};
    }
    // Every concatenated member (tests/io_cli.rs:77-87): inflate until the
    // input is used up; a stream that ends inside a member is an error.
    z_stream z;
    memset(&z, 0, sizeof z);
    if (inflateInit2(&z, 15 + 16) != Z_OK) return "gzip: cannot initialise the decoder";
    z.next_in = const_cast<Bytef *>(data);
    z.avail_in = uInt(len);
    std::vector<uint8_t> buf(1 << 16);
    std::string err;
    for (;;) {
        z.next_out = buf.data();
        z.avail_out = uInt(buf.size());
        const int rc = inflate(&z, Z_NO_FLUSH);
        out.insert(out.end(), buf.data(), buf.data() + (buf.size() - z.avail_out));
        if (rc == Z_STREAM_END) {
            if (z.avail_in == 0) break;
            // Another member follows (only a gzip header may start it).
            if (z.avail_in < 2 || z.next_in[0] != 0x1f || z.next_in[1] != 0x8b) {
                err = "gzip: trailing data after the last member";
                break;
            }
            if (inflateReset(&z) != Z_OK) {
                err = "gzip: cannot restart the decoder";
                break;
            }
            continue;
        }
        if (rc == Z_OK) {
            if (z.avail_in == 0 && z.avail_out != 0) {
                err = "gzip stream is truncated";
                break;
            }
            continue;
        }
        if (rc == Z_BUF_ERROR && z.avail_in == 0) {
            err = "gzip stream is truncated";
            break;
        }
        err = std::string("gzip: ") + (z.msg ? z.msg : "corrupt data");
        break;
    }
    inflateEnd(&z);
    if (!err.empty()) out.clear();
    return err;
}

namespace {

struct Lines {
    const char *p, *end;
    // Next line without its terminator (and a trailing CR); false at the end.
    bool next(std::string &line) {
        if (p >= end) return false;
        const char *q = static_cast<const char *>(memchr(p, '\n', size_t(end - p)));
        const char *e = q ? q : end;
        line.assign(p, e);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        p = q ? q + 1 : end;
        return true;
    }
};

bool blank(const std::string &s) {
    for (char c : s)
        if (!isspace((unsigned char)c)) return false;
    return true;
}

// ">name comment" / "@name comment": the name is the first word, the
// comment the rest with surrounding blanks removed.
bool header(const std::string &line, SeqRecord &r) {
    size_t a = 1;
    while (a < line.size() && isspace((unsigned char)line[a])) a++;
    size_t b = a;
    while (b < line.size() && !isspace((unsigned char)line[b])) b++;
    if (b == a) return false;
    r.name.assign(line, a, b - a);
    size_t c = b;
    while (c < line.size() && isspace((unsigned char)line[c])) c++;
    size_t d = line.size();
    while (d > c && isspace((unsigned char)line[d - 1])) d--;
    r.comment.assign(line, c, d - c);
    return true;
}

void append_bases(const std::string &line, std::vector<uint8_t> &seq) {
    for (char c : line)
        if (!isspace((unsigned char)c)) seq.push_back(uint8_t(c));
}

}  // namespace

std::string parse_sequences(const uint8_t *text, size_t len, std::vector<SeqRecord> &out) {
    out.clear();
    Lines in{reinterpret_cast<const char *>(text), reinterpret_cast<const char *>(text) + len};
    std::string line;
    bool have = in.next(line);
    while (have) {
        if (blank(line)) {
            have = in.next(line);
            continue;
        }
        SeqRecord r;
        if (line[0] == '>') {
            if (!header(line, r)) return "FASTA record has no name";
            // Sequence lines up to the next record; blanks inside lines are dropped.
            while ((have = in.next(line)) && (line.empty() || (line[0] != '>' && line[0] != '@')))
                append_bases(line, r.seq);
            out.push_back(std::move(r));
            continue;
        }
        if (line[0] == '@') {
            if (!header(line, r)) return "FASTQ record has no name";
            // Multi-line sequence up to the '+' separator, then quality lines
            // until they cover the sequence (a quality line may start with
            // '@' or '+').
            bool sep = false;
            while (in.next(line)) {
                if (!line.empty() && line[0] == '+') {
                    sep = true;
                    break;
                }
                append_bases(line, r.seq);
            }
            if (!sep) return "FASTQ record " + r.name + " has no '+' separator";
            while (r.qual.size() < r.seq.size()) {
                if (!in.next(line)) return "FASTQ record " + r.name + " has fewer qualities than bases";
                append_bases(line, r.qual);
            }
            if (r.qual.size() != r.seq.size())
                return "FASTQ record " + r.name + " has more qualities than bases";
            out.push_back(std::move(r));
            have = in.next(line);
            continue;
        }
        return "sequence input is neither FASTA nor FASTQ";
    }
    return {};
}

namespace {

std::vector<std::string> split_tabs(const std::string &line) {
    std::vector<std::string> f;
    size_t a = 0;
    for (;;) {
        size_t b = line.find('\t', a);
        f.emplace_back(line, a, b == std::string::npos ? std::string::npos : b - a);
        if (b == std::string::npos) break;
        a = b + 1;
    }
    return f;
}

struct Segment {
    uint32_t first = 0, last = 0;  // node ids of its first and last base
    std::string seq;
};

}  // namespace

std::string read_gfa(const Params &p, const uint8_t *text, size_t len, Graph &g,
                     std::vector<SeqRecord> &records) {
    records.clear();
    Lines in{reinterpret_cast<const char *>(text), reinterpret_cast<const char *>(text) + len};
    std::unordered_map<std::string, size_t> seg_of;
    std::vector<Segment> segs;
    std::vector<std::vector<std::string>> links, paths;
    std::string line;
    while (in.next(line)) {
        if (blank(line) || line[0] == '#') continue;
        std::vector<std::string> f = split_tabs(line);
        const std::string &kind = f[0];
        if (kind == "H") continue;
        if (kind == "S") {
            if (f.size() < 3 || f[1].empty()) return "GFA segment record is incomplete";
            if (f[2] == "*" || f[2].empty()) return "GFA segment " + f[1] + " has no sequence";
            if (seg_of.count(f[1])) return "GFA segment " + f[1] + " is defined twice";
            for (char c : f[2])
                if (!isalpha((unsigned char)c)) return "GFA segment " + f[1] + " has an invalid base";
            seg_of[f[1]] = segs.size();
            Segment s;
            s.seq = f[2];
            segs.push_back(std::move(s));
        } else if (kind == "L") {
            if (f.size() < 6) return "GFA link record is incomplete";
            links.push_back(std::move(f));
        } else if (kind == "P") {
            if (f.size() < 3) return "GFA path record is incomplete";
            paths.push_back(std::move(f));
        } else {
            return "unsupported GFA record type " + kind;
        }
    }
    // Segments: one node per base in file order, chained within the segment.
    for (Segment &s : segs) {
        uint32_t prev = 0;
        for (size_t k = 0; k < s.seq.size(); k++) {
            const uint32_t id = g.add_node(p.encode(uint8_t(s.seq[k])));
            if (k == 0)
                s.first = id;
            else
                g.link(prev, id, 0);
            prev = id;
        }
        s.last = prev;
    }
    auto find = [&](const std::string &name, size_t &out) {
        auto it = seg_of.find(name);
        if (it == seg_of.end()) return false;
        out = it->second;
        return true;
    };
    // Links in file order: forward orientation and zero overlap only.
    for (const auto &f : links) {
        size_t a, b;
        if (!find(f[1], a) || !find(f[3], b)) return "GFA link references an unknown segment";
        if (f[2] != "+" || f[4] != "+") return "GFA link has an unsupported orientation";
        if (f[5] != "0M") return "GFA link has a nonzero overlap";
        if (a == b) return "GFA link forms a cycle";
        g.link(segs[a].last, segs[b].first, 0);
    }
    // Paths, replayed as reads: add_path with the path's own nodes.
    const int64_t src_base = int64_t(g.node(0).base);
    for (size_t k = 0; k < paths.size(); k++) {
        const auto &f = paths[k];
        SeqRecord r;
        r.name = f[1];
        std::vector<AlignedBase> aln;
        aln.push_back(AlignedBase{0, src_base, -1});
        std::vector<uint32_t> ids;
        if (f[2] != "*") {
            size_t a = 0;
            for (;;) {
                size_t b = f[2].find(',', a);
                std::string step = f[2].substr(a, b == std::string::npos ? std::string::npos : b - a);
                if (step.size() < 2) return "GFA path step is malformed";
                const char o = step.back();
                step.pop_back();
                size_t s;
                if (!find(step, s)) return "GFA path references an unknown segment";
                if (o != '+') return "GFA path has an unsupported orientation";
                for (uint32_t id = segs[s].first;; id++) {
                    ids.push_back(id);
                    if (id == segs[s].last) break;
                }
                r.seq.insert(r.seq.end(), segs[s].seq.begin(), segs[s].seq.end());
                if (b == std::string::npos) break;
                a = b + 1;
            }
        }
        {
            std::vector<uint32_t> sorted = ids;
            std::sort(sorted.begin(), sorted.end());
            if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
                return "GFA path " + r.name + " revisits a node";
        }
        for (uint32_t id : ids) aln.push_back(AlignedBase{int64_t(id), int64_t(g.node(id).base), -1});
        std::string e = g.add_path(k, aln.data(), aln.size());
        if (!e.empty()) return e;
        records.push_back(std::move(r));
    }
    g.weigh_unused_edges();
    for (uint32_t id = 2; id < g.size(); id++) {
        if (g.node(id).in.empty()) g.link(0, id, 1);
        if (g.node(id).out.empty()) g.link(id, 1, 1);
    }
    return g.toposort(false);
}

}  // namespace mpoa
