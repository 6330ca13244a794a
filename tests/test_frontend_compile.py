"""CPU compile checks of the C++ front end: the header builds standalone and, where an MPI
header is present, with the MPI_Comm constructors of the reference interface."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
SRC = r'''
#include "heffte_mi355.hpp"
void use(MPI_Comm comm, std::complex<float>* x, std::complex<float>* y, std::complex<float>* w) {
    heffte::box3d<> inbox({0, 0, 0}, {7, 7, 7}), outbox({0, 0, 0}, {7, 7, 7}, {2, 1, 0});
    heffte::fft3d<heffte::backend::mi355> fft(inbox, outbox, comm);
    fft.forward(x, y, w, heffte::scale::full);
    fft.backward(y, x, w);
    heffte::fft3d_r2c<heffte::backend::mi355> r2c(inbox, inbox.r2c(0), 0, comm);
    heffte::rtransform<heffte::backend::mi355_cos> dct(inbox, inbox, comm);
}
'''


@pytest.mark.skipif(HIPCC is None, reason="hipcc not found")
def test_header_compiles_with_mpi(tmp_path):
    mpi_inc = "/opt/conda/include"
    if not os.path.exists(os.path.join(mpi_inc, "mpi.h")):
        pytest.skip("no mpi.h in this image")
    f = tmp_path / "use.cpp"
    f.write_text(SRC)
    r = subprocess.run([HIPCC, "-std=c++17", "-fsyntax-only", "-x", "hip", "--offload-arch=gfx950",
                        "-I" + os.path.join(ROOT, "include"), "-I" + mpi_inc, str(f)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
