"""The host-side code under AddressSanitizer + UndefinedBehaviorSanitizer: the
CPU index builder, the index file reader / writer and the text formatter
(hpg-libs_amd/csrc/check/host_check.cpp: every truncation of index files,
flipped bytes with the checksum kept and recomputed, random records to
format), the importer of the reference's index directories (truncated files
and flipped bytes in each file of a directory the reference's own writers
made), and the oracle's search functions on random texts and reads
(oracle/check_oracle.c).  A sanitizer report fails the run.  CPU only."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import fmoracle as fo
from conftest import ROOT
from hpgfm import synth

ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=66",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
HOST = os.path.join(ROOT, "hpg-libs_amd", "csrc", "_build", "host_check")
ORACLE = os.path.join(ROOT, "oracle", "_check", "check_oracle")


@pytest.fixture(scope="module")
def built():
    for d in (os.path.join(ROOT, "hpg-libs_amd", "csrc"), os.path.join(ROOT, "oracle")):
        r = subprocess.run(["make", "-C", d, "check"], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
    return True


def run(args, timeout=600):
    r = subprocess.run(args, capture_output=True, text=True, env=ENV, timeout=timeout)
    assert r.returncode == 0, (args, r.returncode, r.stdout[-2000:], r.stderr[-4000:])
    return r.stdout


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_host_code_selftest(built, tmp_path, seed):
    run([HOST, "selftest", str(seed), str(tmp_path)])


def test_oracle_under_sanitizers(built):
    for seed in (1, 7):
        assert "ok" in run([ORACLE, str(seed)])


@pytest.mark.skipif(not fo.ref_available(), reason="oracle/_ref not built")
def test_importer_on_damaged_directories(built, tmp_path):
    g, _, _ = synth.genome(3000, seed=12)
    work = str(tmp_path / "work")
    os.makedirs(work)
    r = fo.RefCore(g, ratio=8, workdir=work)
    src = str(tmp_path / "ref")
    r.save_dir(src, [0, 0], [1499, 1499])
    r.close()
    assert run([HOST, "import", src]).strip() == "0"
    rng = np.random.default_rng(13)
    statuses = set()
    for name in sorted(os.listdir(src)):
        data = open(os.path.join(src, name), "rb").read()
        cuts = sorted(set([0, 1, len(data) // 2, max(len(data) - 1, 0)] +
                          list(rng.integers(0, max(len(data), 1), 4))))
        for t, cut in enumerate(cuts):
            d = str(tmp_path / f"cut_{name}_{t}")
            shutil.copytree(src, d)
            open(os.path.join(d, name), "wb").write(data[:cut])
            statuses.add(int(run([HOST, "import", d])))
            shutil.rmtree(d)
        for t in range(6):
            if not data:
                break
            b = bytearray(data)
            b[int(rng.integers(0, len(b)))] ^= 1 << int(rng.integers(0, 8))
            d = str(tmp_path / f"flip_{name}_{t}")
            shutil.copytree(src, d)
            open(os.path.join(d, name), "wb").write(bytes(b))
            statuses.add(int(run([HOST, "import", d])))
            shutil.rmtree(d)
    assert statuses <= {0, -3, -4}
    assert -4 in statuses


def test_loader_on_files_written_by_the_library(built, tmp_path):
    import hpgfm
    g, _, _ = synth.genome(20_000, seed=14, repeat_frac=0.2, fam_len=(100, 600))
    ix = hpgfm.Index.build(g, 16)
    ix.set_karyotype([0, 5], [9999, 10004])
    p = str(tmp_path / "i.fmi")
    ix.save(p)
    assert run([HOST, "load", p]).strip() == "0"
    data = open(p, "rb").read()
    for cut in (0, 7, 31, 200, len(data) // 3, len(data) - 9, len(data) - 1):
        q = str(tmp_path / f"c{cut}.fmi")
        open(q, "wb").write(data[:cut])
        assert run([HOST, "load", q]).strip() == "-4"
