"""GPU parity on the measured workloads' models and through the reference's executable path.

- The in-core workload (50 Mbp i.i.d. genome, 100 bp reads at 30x, 50 % reverse-complemented,
  0.5 % substitutions, -o 63) on a 2 % sample of the same coverage: output.bin byte for byte and
  every round's live counts (S, P before and after removal, E) against the oracle, through the
  kept-read and the FASTQ entry points.
- The empanada CLI (main.cu:42-71: the block loop of read_block with its seek-back, then
  build_graph writing output.bin) run as a child process with a block size that cuts the FASTQ
  into dozens of blocks, against the oracle's reader and overlap path on the same file."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
from empanada_amd import gen, native

pytestmark = pytest.mark.gpu

COUNTS = ("s", "p", "edges", "s_after", "p_after")


def _counts(rounds):
    return [tuple(x[k] for k in COUNTS) for x in rounds]


def test_incore_workload_sample_vs_oracle():
    r, o = gen.workload_reads("incore", 0.02)
    n, L = r.shape
    assert (n, L, o) == (300_000, 100, 63)
    rc, want, osum, orounds = ol.run_reads(r.tobytes(), n, L, min_ovlp=o)
    assert rc == 0 and osum["n_rounds"] == 37 and osum["n_edges"] > 0
    with native.Context(min_ovlp=o) as c:
        c.add_reads(r)
        got = c.run()
        summ, rounds = c.summary(), c.rounds()
    assert got == want
    assert summ["n_dedup"] == osum["n_dedup"] and summ["n_records"] == osum["n_records"] == 37
    assert _counts(rounds) == _counts(orounds)
    fq = gen.fastq_bytes(r)
    with native.Context(min_ovlp=o) as c:
        c.add_fastq(fq, block=16 << 20)
        assert c.run() == want


def test_fastq_file_and_device_entry_points(tmp_path):
    """emp_add_fastq_file (the file streamed into HBM in 64 MiB chunks while blocks are parsed) and
    emp_add_fastq_device (blocks already in HBM) give the oracle's output at the default block of
    64 MiB, which straddles the chunks of a 3 % in-core sample (96 MB of FASTQ), and at blocks of
    100 KiB; a block shorter than one record is the format error of the oracle's reader."""
    from empanada_amd import hipmem
    r, o = gen.workload_reads("incore", 0.03)
    fq = tmp_path / "in.fastq"
    gen.write_fastq(str(fq), r)
    data = fq.read_bytes()
    assert len(data) > 64 << 20
    rc, want, _, _ = ol.run_fastq(data, min_ovlp=o)
    assert rc == 0
    for map_mem in (1 << 30, 1600 << 10):
        assert ol.run_fastq(data, min_ovlp=o, map_mem=map_mem)[1] == want
        with native.Context(min_ovlp=o, map_mem=map_mem) as c:
            c.add_fastq_file(str(fq))
            assert c.run() == want, map_mem
    d = hipmem.DeviceBuffer.from_array(np.frombuffer(data, np.uint8))
    for map_mem, pieces in ((1 << 30, 1), (1 << 30, 4), (16 * ((100 << 10) + 3), 3)):
        with native.Context(min_ovlp=o, map_mem=map_mem) as c:
            c.add_fastq_device(d.address, len(data), pieces=pieces)
            assert c.run() == want, (map_mem, pieces)
    d.free()
    assert ol.run_fastq(data, min_ovlp=o, map_mem=16 * 150)[0] == -2
    with native.Context(min_ovlp=o, map_mem=16 * 150) as c:
        with pytest.raises(native.EmpError) as e:
            c.add_fastq_file(str(fq))
    assert e.value.code == -3


def test_large_pageable_reads_stream_through_pinned_stages():
    """3,600,003 x 150 bp reads in pageable memory (540 MB, past the size where emp_add_reads
    streams them through three pinned stages, packed to 2-bit codes on the host and expanded on
    the device: batches of 1,789,568 reads, the third one short) against the oracle; the last 3
    reads (off a multiple of 16) are staged. Every 7th read in lower case gives the same output
    (lower case codes as upper case), also from pinned host memory (an array this large is packed
    through the stages as well). One byte other than ACGTacgt in the second batch: that batch goes
    over unpacked (pinned: straight from the caller's buffer) and the map kernel's check reports
    EMP_ERR_BASE."""
    from empanada_amd import hipmem
    g = gen.genome(3 * 10**7, 81)
    r = gen.reads(g, 3_600_003, 150, 0.5, 0.005, 82)
    rc, want, osum, _ = ol.run_reads(r.tobytes(), r.shape[0], 150, min_ovlp=75)
    assert rc == 0 and osum["n_edges"] > 0
    with native.Context(min_ovlp=75) as c:
        c.add_reads(r)
        assert c.run() == want
    r[::7] |= 0x20
    h = hipmem.HostBuffer(r.nbytes)
    hr = h.array(np.uint8, r.shape)
    hr[:] = r
    for src in (r, hr):
        with native.Context(min_ovlp=75) as c:
            c.add_reads(src)
            assert c.run() == want
    r[3_000_001, 77] = ord("N")
    hr[3_000_001, 77] = ord("N")
    for src in (r, hr):
        with native.Context(min_ovlp=75) as c:
            with pytest.raises(native.EmpError) as e:
                c.add_reads(src)
                c.run()
        assert e.value.code == -5
    h.free()


def test_cli_block_loop_vs_oracle(tmp_path):
    g = gen.genome(3 * 10**5, 51)
    r = gen.reads(g, 30000, 100, 0.5, 0.005, 52)
    fq = tmp_path / "in.fastq"
    gen.write_fastq(str(fq), r)
    data = fq.read_bytes()
    out = tmp_path / "output.bin"
    # -m 1600k: blocks of 100 KiB, about 65 of them, each resumed after its last complete record
    p = subprocess.run([native.CLI_PATH, str(fq), str(out), "-o", "40", "-m", "1600k"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert len(data) > 60 * (100 << 10)
    rc, want, osum, _ = ol.run_fastq(data, min_ovlp=40, map_mem=1600 << 10)
    assert rc == 0 and osum["n_edges"] > 0
    assert out.read_bytes() == want
    assert "edges=%d" % osum["n_edges"] in p.stderr


def test_cli_errors(tmp_path):
    missing = subprocess.run([native.CLI_PATH, str(tmp_path / "none.fastq"), str(tmp_path / "o.bin")],
                             capture_output=True, text=True, timeout=60)
    assert missing.returncode == 1 and "cannot open" in missing.stderr
    bad = tmp_path / "bad.fastq"
    bad.write_bytes(b"@r0\nACGT\n+\nIIII\n@r1\nACG\n+\nIII\n")
    p = subprocess.run([native.CLI_PATH, str(bad), str(tmp_path / "o.bin")], capture_output=True, text=True,
                       timeout=60)
    assert p.returncode == 1 and "error -4" in p.stderr
    assert not os.path.exists(tmp_path / "o.bin") or os.path.getsize(tmp_path / "o.bin") == 0


@pytest.mark.parametrize("direct", ["1", "0"])
def test_full_incore_workload_to_a_new_output_file(direct, tmp_path, monkeypatch):
    """The whole in-core workload (15 M reads; records of up to 22 MB, 143 MB of output.bin)
    written to a new file: with direct writes allowed, the records that arrive while the buffered
    writes lag behind have a share of their 2 MB chunks written with O_DIRECT beside them; the
    file equals emp_run's records byte for byte either way, and so does a second run into a new
    file on the same context (the direct writes' AIO context and buffers are reused)."""
    monkeypatch.setenv("EMPANADA_DIRECT_WRITES", direct)
    r, o = gen.workload_reads("incore", 1.0)
    with native.Context(min_ovlp=o) as c:
        c.add_reads(r)
        want = c.run()
        assert len(want) > 100 << 20
        for k in range(2):
            out = tmp_path / ("output%d.bin" % k)
            c.reset()
            c.add_reads(r)
            c.run_to_file(str(out))
            assert out.stat().st_size == len(want)
            assert out.read_bytes() == want
            out.unlink()
