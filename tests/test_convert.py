"""output.bin -> GFA / CSV converter (empanada_amd/bin/emp-convert), on the CPU.

The GFA check is semantic: every L line must name two oriented reads whose sequences really
overlap by the stated length (suffix of From in its orientation = prefix of To in its
orientation), and the lines must be the oracle's edges in file order."""
import os
import random
import struct
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONV = os.path.join(ROOT, "empanada_amd", "bin", "emp-convert")
COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def rc(s):
    return "".join(COMP[c] for c in reversed(s))


def records(blob):
    out, off = [], 0
    while off < len(blob):
        n, ovlp = struct.unpack_from("<Qi", blob, off)
        off += 12
        fr = np.frombuffer(blob, dtype=np.uint32, count=n, offset=off)
        off += 4 * n
        to = np.frombuffer(blob, dtype=np.uint32, count=n, offset=off)
        off += 4 * n
        out.append((ovlp, fr, to))
    return out


@pytest.fixture(scope="module")
def job(tmp_path_factory):
    rng = random.Random(3)
    genome = "".join(rng.choice("ACGT") for _ in range(400))
    reads = []
    for _ in range(60):
        p = rng.randrange(0, 400 - 40)
        s = genome[p:p + 40]
        reads.append(rc(s) if rng.random() < 0.5 else s)
    data = "".join(reads).encode()
    r, out, _, _ = ol.run_reads(data, len(reads), 40, min_ovlp=12)
    assert r == 0
    d = tmp_path_factory.mktemp("conv")
    path = d / "output.bin"
    path.write_bytes(out)
    return reads, out, path, d


def test_gfa_lines_are_real_overlaps_in_file_order(job):
    reads, out, path, d = job
    gfa = d / "out.gfa"
    subprocess.run([CONV, str(path), str(gfa), "--gfa", "--reads", str(len(reads)), "--len", "40"], check=True,
                   capture_output=True)
    lines = gfa.read_text().splitlines()
    assert lines[0] == "H\tVN:Z:1.0"
    segs = [ln for ln in lines if ln.startswith("S\t")]
    assert segs == ["S\t%d\t*\tLN:i:40" % k for k in range(len(reads))]
    links = [ln.split("\t") for ln in lines if ln.startswith("L\t")]
    want = [(int(f), int(t), ov) for ov, fr, to in records(out) for f, t in zip(fr, to)]
    assert len(links) == len(want) and len(want) > 0
    for (_, u, uo, v, vo, cig), (f, t, ov) in zip(links, want):
        assert (int(u), uo, int(v), vo, cig) == (f // 2, "-" if f & 1 else "+", t // 2, "-" if t & 1 else "+",
                                                  "%dM" % ov)
        su = reads[int(u)] if uo == "+" else rc(reads[int(u)])
        sv = reads[int(v)] if vo == "+" else rc(reads[int(v)])
        assert su[-ov:] == sv[:ov]


def test_csv_lines(job):
    reads, out, path, d = job
    csv = d / "out.csv"
    subprocess.run([CONV, str(path), str(csv), "--csv"], check=True, capture_output=True)
    want = ["%d,%d,%d" % (f, t, ov) for ov, fr, to in records(out) for f, t in zip(fr, to)]
    assert csv.read_text().splitlines() == want


def test_truncated_input_fails(job, tmp_path):
    _, out, _, _ = job
    bad = tmp_path / "bad.bin"
    bad.write_bytes(out[: len(out) - 3])
    r = subprocess.run([CONV, str(bad), str(tmp_path / "x.gfa"), "--gfa"], capture_output=True)
    assert r.returncode == 1
    r = subprocess.run([CONV, str(bad)], capture_output=True)
    assert r.returncode == 2
