"""CPU checks of the C ABI library: it loads and exports every symbol the headers declare, and the
host-only paths (argument checks, FASTQ block parsing errors raised before any device work) behave.
No compute call needs a GPU here."""
import ctypes
import os
import re
import subprocess

import pytest

from empanada_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(emp_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_headers_match_binding_list():
    decl = declared_symbols()
    assert decl == set(native.EXPORTS), sorted(decl ^ set(native.EXPORTS))


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert declared_symbols() <= exported


def test_version_and_defaults():
    lib = native.lib()
    assert b"gfx950" in lib.emp_version()
    o = native.Opts()
    lib.emp_opts_default(ctypes.byref(o))
    # config.h:124-128 defaults
    assert (o.min_ovlp, o.map_mem, o.reduce_mem, o.base_blocksz, o.kv_chunksz) == (31, 1 << 30, 1 << 30, 32 << 20,
                                                                                   64 << 20)


def test_null_arguments_are_rejected_without_a_device():
    lib = native.lib()
    assert lib.emp_ctx_create(None, None) == -1
    assert lib.emp_add_reads(None, None, 0, 100) == -1
    assert lib.emp_run(None, native.SINK(lambda *a: 0), None) == -1
    assert lib.emp_stage_roll(0, None, None, 5, 3, 10, None) == -1


def test_cli_usage_without_arguments():
    r = subprocess.run([native.CLI_PATH], capture_output=True, text=True)
    assert r.returncode == 2 and "Usage" in r.stdout


def test_cli_rejects_bad_option_and_unit():
    r = subprocess.run([native.CLI_PATH, "a.fq", "b.bin", "-x", "3"], capture_output=True, text=True)
    assert r.returncode == 2 and "Invalid option" in r.stdout
    r = subprocess.run([native.CLI_PATH, "a.fq", "b.bin", "-m", "3q"], capture_output=True, text=True)
    assert r.returncode == 2


def test_repeat_copies_never_overlap():
    """The 3.1 Gbp workload's repeat families: copies of both sizes (6 kbp first, then 300 bp) take
    disjoint slots of one occupancy map, and every copy's slot is marked."""
    import numpy as np

    from empanada_amd import gen
    g = gen.genome(3_000_000, 1200)
    used = np.zeros(g.size // 300, dtype=np.uint8)
    f6 = gen.add_repeats(g, 1208, 0.05, 6000, 16, 0.01, used, 300)
    after6 = used.copy()
    f3 = gen.add_repeats(g, 1207, 0.05, 300, 16, 0.01, used, 300)
    assert f6 >= 1 and f3 >= 1
    assert after6.sum() == 20 * 16 * f6                       # whole 6 kbp slots, 20 cells each
    assert used.sum() == after6.sum() + 16 * f3               # 300 bp copies only on free cells
    assert (after6.reshape(-1, 20).sum(axis=1) % 20 == 0).all()  # 6 kbp copies on their own grid
