"""The Taskflow/Threadpool drop-in used as a caller of the reference would use it: a C++ program with the
2d_cholesky.cpp PTG written against include/tasktorrent_amd/taskflow.hpp, compiled with g++ and linked
against the C-ABI library. Without a device its compiled graph must equal the oracle's DepsLogger edges,
and a bound task (set_binding true) must be refused, not ignored."""
import subprocess

import pytest

import dropin_build


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    return dropin_build.build(tmp_path_factory.mktemp("dropin"))


@pytest.mark.parametrize("nb,prio", [(1, 2), (5, 2), (7, 0)])
def test_dropin_graph_equals_oracle_deps(exe, oracle, tmp_path, nb, prio):
    r = subprocess.run([str(exe), "dry", str(nb), str(prio)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    ours = [tuple(line.split(",")) for line in r.stdout.splitlines() if line]
    A = oracle.spd_bbt(nb * 8, 2)
    deps = tmp_path / "deps.csv"
    oracle.cholesky(A, 8, n_threads=2, prio_kind=prio, deps_path=str(deps))
    theirs = {tuple(line.split(",")) for line in deps.read_text().splitlines() if line}
    assert len(ours) == len(set(ours))
    assert set(ours) == theirs


def test_dropin_binding_is_refused(exe):
    r = subprocess.run([str(exe), "binding", "4"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert int(r.stdout.strip()) == 1  # TTAD_ERR_INVALID
    assert "binding = true for TRSM_" in r.stderr


@pytest.mark.gpu
def test_dropin_factor_on_gpu(exe, oracle, tmp_path):
    import numpy as np
    n, b = 1280, 256
    A = oracle.spd_bbt(n, 77)
    src, dst = tmp_path / "A.bin", tmp_path / "L.bin"
    np.asfortranarray(A).T.tofile(src)  # column-major bytes
    print(f"drop-in GPU run n={n} b={b}", flush=True)
    r = subprocess.run([str(exe), "gpu", str(n), str(b), str(src), str(dst)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    L = np.fromfile(dst, dtype=np.float64).reshape(n, n).T
    Lref, _ = oracle.cholesky(A, b, n_threads=2)
    assert np.linalg.norm(L - Lref) / np.linalg.norm(Lref) < 1e-12
