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


TRAITS_SRC = r'''
#include "heffte_mi355.hpp"
#include <type_traits>
using namespace heffte;
// the tag alone selects the library's specializations and traits
static_assert(std::is_base_of<mi355_detail::fft3d_impl<backend::mi355, int>, fft3d<backend::mi355>>::value, "fft3d");
static_assert(std::is_base_of<mi355_detail::fft3d_impl<backend::mi355_cos, long long>,
                              rtransform<backend::mi355_cos, long long>>::value, "rtransform");
static_assert(std::is_base_of<mi355_detail::fft3d_r2c_impl<int>, fft3d_r2c<backend::mi355>>::value, "fft3d_r2c");
static_assert(backend::is_enabled<backend::mi355>::value && backend::is_enabled<backend::mi355_cos1>::value, "enabled");
static_assert(std::is_same<backend::buffer_traits<backend::mi355>::location, tag::gpu>::value, "gpu location");
static_assert(backend::uses_fft_types<backend::mi355>::value && !backend::uses_fft_types<backend::mi355_sin>::value,
              "fft types");
static_assert(!default_plan_options<backend::mi355>::use_reorder, "no reorder pass");
int main() {
    plan_options o = default_options<backend::mi355>();
    return (o.use_reorder || backend::name<backend::mi355>() != "mi355") ? 1 : 0;
}
'''


@pytest.mark.skipif(HIPCC is None, reason="hipcc not found")
def test_backend_tag_traits_and_specializations(tmp_path):
    f = tmp_path / "traits.cpp"
    f.write_text(TRAITS_SRC)
    r = subprocess.run([HIPCC, "-std=c++17", "-fsyntax-only", "-x", "hip", "--offload-arch=gfx950",
                        "-DHEFFTE_MI355_NO_MPI", "-I" + os.path.join(ROOT, "include"), str(f)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(HIPCC is None, reason="hipcc not found")
def test_plugin_header_needs_heffte_declarations(tmp_path):
    """heffte_backend_mi355.h adds to heFFTe's namespace; alone it stops with a clear error."""
    f = tmp_path / "alone.cpp"
    f.write_text('#include "heffte_backend_mi355.h"\n')
    r = subprocess.run([HIPCC, "-std=c++17", "-fsyntax-only", "-x", "hip", "--offload-arch=gfx950",
                        "-I" + os.path.join(ROOT, "include"), str(f)], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "include heffte.h" in r.stderr
