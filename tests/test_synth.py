"""The workload generators (CPU): seeded and reproducible, with the stated
shapes of each configuration."""
import numpy as np

from minipoa_hip import Params, synth, tune


def test_protein_families_shape_and_matrix():
    p, jo, ro, b = synth.protein_families(nfam=40, members=16)
    assert len(jo) == 41 and int(jo[-1]) == 40 * 16
    n = len(synth.PROTEIN)
    m = np.array(p.matrix).reshape(n, n)
    assert (m == m.T).all()
    assert ((np.diag(m) >= 4) & (np.diag(m) <= 11)).all()
    off = m[~np.eye(n, dtype=bool)]
    assert off.min() >= -4 and off.max() <= 3
    assert p.f == 0 and bytes(p.alphabet) == synth.PROTEIN
    assert set(np.unique(b).tolist()) <= set(synth.PROTEIN)
    again = synth.protein_families(nfam=40, members=16)
    assert (again[3] == b).all() and again[0].matrix == p.matrix


def test_diverge_event_counts():
    rng = np.random.Generator(np.random.PCG64(1))
    root = rng.integers(0, 20, size=1000)
    out = synth.diverge(np.random.Generator(np.random.PCG64(2)), root, 20, 0.0)
    assert (out == root).all()
    subs = synth.diverge(np.random.Generator(np.random.PCG64(3)), root, 20, 0.3, sub=1.0, ins=0.0)
    assert len(subs) == 1000 and int((subs != root).sum()) == 300


def test_serving_stream_arrivals_and_incremental_jobs():
    s = synth.serving_stream(njobs=200, rate_per_s=10.0)
    t = [j["t"] for j in s]
    assert all(a < b for a, b in zip(t, t[1:]))
    assert 10 < len(s) / t[-1] * 2 and len(s) / t[-1] < 20
    incr = [j for j in s if j["base"]]
    assert 5 <= len(incr) <= 40
    for j in incr:
        assert len(j["base"]) == 50 and len(j["reads"]) == 20
    for j in s:
        if not j["base"]:
            assert 5 <= len(j["reads"]) <= 64
        L = len(j["reads"][0])
        assert 200 < L < 24000


def test_tune_calibration_rows_match_their_bucket():
    for width in (384, 1024, 4096):
        p, jo, ro, b = tune.calibration(width, cells=1e8)
        L = int(ro[1] - ro[0])
        m = L + 2
        band = p.b + m // p.f
        assert abs(2 * band + 16 - width) <= width // 4 + 32, (width, band)
        assert p.adaptive and len(jo) >= 17


def test_matrix_text_round_trips_through_load_matrix(tmp_path):
    import minipoa_hip as mh

    p, *_ = synth.protein_families(nfam=2, members=2)
    f = tmp_path / "p.mtx"
    f.write_text(synth.matrix_text(p.alphabet, p.matrix))
    q = mh.load_matrix(Params(f=0), f)
    assert bytes(q.alphabet) == bytes(p.alphabet) and q.matrix == p.matrix
