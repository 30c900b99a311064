"""The Fortran host (module mw3_capi, ISO_C_BINDING) drives libmw3 on the GPU
and gets the same results as the ctypes host on the same case."""
import os
import subprocess

import numpy as np
import pytest

from madwave3_amd.lib import MW3_FP64, MW3_MIXED, Propagator, read_fdrive_result, write_case_file
from tests.cases import case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDRIVE = os.path.join(ROOT, "madwave3_amd", "fortran", "mw3_fdrive")


@pytest.mark.gpu
@pytest.mark.parametrize("name,precision", [("coriol", MW3_FP64), ("coriol", MW3_MIXED),
                                            ("prodJ", MW3_FP64)])
def test_fortran_host_matches_python_host(tmp_path, name, precision):
    assert os.path.exists(FDRIVE), "mw3_fdrive is not built"
    c = case(name)
    nsteps = 120
    cf = str(tmp_path / "case.bin")
    rf = str(tmp_path / "res.bin")
    write_case_file(cf, c, c.rpaq0, nsteps, precision)
    print("fortran host: case %s, %d steps, precision %d" % (name, nsteps, precision), flush=True)
    r = subprocess.run([FDRIVE, cf, rf], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    norms_f, P_f, z_f, zp_f = read_fdrive_result(rf)

    g = Propagator(c, precision)
    g.begin(c.rpaq0)
    norms_p, _ = g.steps(nsteps)
    P_p = g.flux_result()[0]
    z_p = g.zs2()
    zp_p = g.zs2prod() if c.iprod == 2 else np.zeros(0)
    g.close()
    z_p = np.stack([z_p.real, z_p.imag], axis=-1).ravel()
    zp_p = np.stack([zp_p.real, zp_p.imag], axis=-1).ravel()
    # same library, same inputs, deterministic kernels: identical results
    assert np.array_equal(norms_f, norms_p)
    assert np.array_equal(P_f, P_p)
    assert np.array_equal(z_f, z_p)
    assert zp_f.size == zp_p.size and np.array_equal(zp_f, zp_p)
