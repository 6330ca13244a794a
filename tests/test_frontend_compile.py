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


C_BINDING_SRC = r'''
#include "heffte_mi355.hpp"
#include <complex>
#define Heffte_BACKEND_MI355 12
// The branches a maintainer adds to heFFTe's C binding for the new id (src/heffte_c.cpp
// dispatch sites), over the plan record that binding keeps (backend id, r2c flag, type-erased
// fft object).  Each branch constructs, calls and destroys through the backend tag alone.
struct plan_record { int backend_type; int use_r2c; void *fft; };
static bool backend_valid(int backend) { return backend == Heffte_BACKEND_MI355; }
static heffte::plan_options options_for(int backend) {
    if (backend == Heffte_BACKEND_MI355) return heffte::default_options<heffte::backend::mi355>();
    return heffte::plan_options(false, heffte::reshape_algorithm::alltoallv, true);
}
static int plan_create(int backend, heffte::box3d<> const &inbox, heffte::box3d<> const &outbox, int r2c_direction,
                       MPI_Comm comm, plan_record *plan) {
    if (!backend_valid(backend)) return 1;
    heffte::plan_options opts = options_for(backend);
    plan->backend_type = backend;
    plan->use_r2c = r2c_direction >= 0;
    try {
        if (plan->use_r2c)
            plan->fft = reinterpret_cast<void *>(
                new heffte::fft3d_r2c<heffte::backend::mi355>(inbox, outbox, r2c_direction, comm, opts));
        else
            plan->fft = reinterpret_cast<void *>(new heffte::fft3d<heffte::backend::mi355>(inbox, outbox, comm, opts));
    } catch (std::invalid_argument &) {
        return 2;
    }
    return 0;
}
static int forward_c2c(plan_record const *plan, std::complex<float> const *in, std::complex<float> *out,
                       std::complex<float> *workspace, heffte::scale s) {
    if (plan->backend_type != Heffte_BACKEND_MI355) return 1;
    if (plan->use_r2c) return 3;
    reinterpret_cast<heffte::fft3d<heffte::backend::mi355> const *>(plan->fft)->forward(in, out, workspace, s);
    return 0;
}
static int forward_r2c(plan_record const *plan, float const *in, std::complex<float> *out,
                       std::complex<float> *workspace, heffte::scale s) {
    if (plan->backend_type != Heffte_BACKEND_MI355 || !plan->use_r2c) return 1;
    reinterpret_cast<heffte::fft3d_r2c<heffte::backend::mi355> const *>(plan->fft)->forward(in, out, workspace, s);
    return 0;
}
static long long size_workspace(plan_record const *plan) {
    if (plan->use_r2c)
        return reinterpret_cast<heffte::fft3d_r2c<heffte::backend::mi355> const *>(plan->fft)->size_workspace();
    return reinterpret_cast<heffte::fft3d<heffte::backend::mi355> const *>(plan->fft)->size_workspace();
}
static void plan_destroy(plan_record *plan) {
    if (plan->backend_type != Heffte_BACKEND_MI355) return;
    if (plan->use_r2c) delete reinterpret_cast<heffte::fft3d_r2c<heffte::backend::mi355> *>(plan->fft);
    else delete reinterpret_cast<heffte::fft3d<heffte::backend::mi355> *>(plan->fft);
}
int use(MPI_Comm comm, float *r, std::complex<float> *x, std::complex<float> *w) {
    heffte::box3d<> box({0, 0, 0}, {7, 7, 7});
    plan_record c2c{}, r2c{};
    int st = plan_create(Heffte_BACKEND_MI355, box, box, -1, comm, &c2c);
    st += plan_create(Heffte_BACKEND_MI355, box, box.r2c(0), 0, comm, &r2c);
    st += forward_c2c(&c2c, x, x, w, heffte::scale::none) + forward_r2c(&r2c, r, x, w, heffte::scale::full);
    st += (int)(size_workspace(&c2c) + size_workspace(&r2c));
    plan_destroy(&c2c);
    plan_destroy(&r2c);
    return st;
}
'''


@pytest.mark.skipif(HIPCC is None, reason="hipcc not found")
def test_c_binding_branches_compile(tmp_path):
    """The Heffte_BACKEND_MI355 branches of heFFTe's C binding (as written out for maintainers)
    compile against the tag's fft3d / fft3d_r2c specializations with MPI_Comm constructors."""
    mpi_inc = "/opt/conda/include"
    if not os.path.exists(os.path.join(mpi_inc, "mpi.h")):
        pytest.skip("no mpi.h in this image")
    f = tmp_path / "cbind.cpp"
    f.write_text(C_BINDING_SRC)
    r = subprocess.run([HIPCC, "-std=c++17", "-fsyntax-only", "-x", "hip", "--offload-arch=gfx950",
                        "-I" + os.path.join(ROOT, "include"), "-I" + mpi_inc, str(f)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
