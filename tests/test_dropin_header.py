"""The drop-in C++ header compiles for a caller written against the reference's surface
(inc/cudecompose.h:23-52) and links against libtdx.so with no GPU headers.

Each wrapper is bound to a function pointer spelled with the reference's parameter list,
so a changed type or argument count fails to compile.  The calls sit behind a branch
the program never takes (argc > 99): the executable loads the library and exits without
touching a GPU, so this runs on the CPU."""
import os
import shutil
import subprocess

import pytest

import tdx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CALLER = r"""
#include "cudecompose.h"

/* the reference's prototypes, inc/cudecompose.h:23-52, as pointer types */
static void (*p_hadmard)(dt *, dt *, dt *, long, long) = hadmard;
static void (*p_hadmard3)(dt *, dt *, dt *, dt *, long, long) = hadmard3;
static void (*p_kr)(dt *, dt *, dt *, long, long, long) = kr;
static void (*p_tensor_product)(dt *, dt *, dt *, long, long, long, long) = tensor_product;
static void (*p_printTensor)(dt *, long, long, long) = printTensor;
static void (*p_f2h)(dt *, half *, long) = f2h;
static void (*p_ttm)(dt *, dt *, dt *, long, long, long, long, long) = ttm;
static void (*p_mttkrp)(dt *, dt *, dt *, dt *, long, long, long, long, long) = mttkrp;
static void (*p_t2m)(dt *, dt *, long, long, long, long) = t2m;
static void (*p_tminust)(dt *, dt *, long) = tminust;
static void (*p_rse)(dt *, dt *, long, dt *) = rse;
static void (*p_gencp)(dt *, dt *, dt *, dt *, long, long, long, long) = gencp;
static void (*p_gentucker)(dt *, dt *, dt *, dt *, dt *, long, long, long, long, long, long) = gentucker;
static void (*p_gencp4)(dt *, dt *, dt *, dt *, dt *, long, long, long, long, long) = gencp4;
static void (*p_gentucker4)(dt *, dt *, dt *, dt *, dt *, dt *, long, long, long, long, long, long, long,
                            long) = gentucker4;
static void (*p_cp_als)(dt *, dt *, dt *, dt *, long, long, long, long) = cp_als;
static void (*p_cp_als4)(dt *, dt *, dt *, dt *, dt *, long, long, long, long, long) = cp_als4;
static void (*p_cp_tensorcore)(dt *, dt *, dt *, dt *, long, long, long, long) = cp_tensorcore;
static void (*p_tucker_hosvd)(dt *, dt *, dt *, dt *, dt *, long, long, long, long, long, long) = tucker_hosvd;
static void (*p_tucker_tensorcore)(dt *, dt *, dt *, dt *, dt *, long, long, long, long, long,
                                   long) = tucker_tensorcore;
static void (*p_tucker_hosvd4)(dt *, dt *, dt *, dt *, dt *, dt *, long, long, long, long, long, long, long,
                               long) = tucker_hosvd4;
static void (*p_initMat)(dt *, long) = initMat;

int main(int argc, char **argv) {
    (void)argv;
    if (argc > 99) {  /* never taken: the calls only make the link resolve every entry point */
        dt *x = 0, e = 0;
        half *h = 0;
        p_hadmard(x, x, x, 2, 2);
        p_hadmard3(x, x, x, x, 2, 2);
        p_kr(x, x, x, 2, 2, 2);
        p_tensor_product(x, x, x, 2, 2, 2, 2);
        p_printTensor(x, 2, 2, 2);
        p_f2h(x, h, 8);
        p_ttm(x, x, x, 1, 2, 2, 2, 2);
        p_mttkrp(x, x, x, x, 1, 2, 2, 2, 2);
        p_t2m(x, x, 1, 2, 2, 2);
        p_tminust(x, x, 4);
        p_rse(x, x, 8, &e);
        p_gencp(x, x, x, x, 2, 2, 2, 2);
        p_gentucker(x, x, x, x, x, 2, 2, 2, 2, 2, 2);
        p_gencp4(x, x, x, x, x, 2, 2, 2, 2, 2);
        p_gentucker4(x, x, x, x, x, x, 2, 2, 2, 2, 2, 2, 2, 2);
        p_cp_als(x, x, x, x, 2, 2, 2, 2);
        p_cp_als4(x, x, x, x, x, 2, 2, 2, 2, 2);
        p_cp_tensorcore(x, x, x, x, 2, 2, 2, 2);
        p_tucker_hosvd(x, x, x, x, x, 2, 2, 2, 2, 2, 2);
        p_tucker_tensorcore(x, x, x, x, x, 2, 2, 2, 2, 2, 2);
        p_tucker_hosvd4(x, x, x, x, x, x, 2, 2, 2, 2, 2, 2, 2, 2);
        p_initMat(x, 4);
        return tdx_last_status();
    }
    printf("dropin-ok\n");
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_reference_caller_compiles_links_and_loads(tmp_path):
    src = tmp_path / "caller.cpp"
    exe = tmp_path / "caller"
    src.write_text(CALLER)
    libdir = os.path.dirname(tdx.LIB_PATH)
    cmd = ["g++", "-std=c++11", "-Wall", "-Werror", "-Wno-unused-variable", "-I", os.path.join(ROOT, "include"),
           str(src), "-o", str(exe), "-L", libdir, "-ltdx", "-Wl,-rpath," + libdir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "dropin-ok"


LIMITS = r"""
#include <string.h>
#include "cudecompose.h"

/* Inputs beyond the documented limits (rank > 64 for the CP drivers, a mode longer than 128
 * for the Tucker drivers) are rejected by argument validation before any device work, so the
 * pointers are never dereferenced and this runs without a GPU. */
int main() {
    static float buf[16];
    dt *x = buf;
    int bad = 0;
    cp_als(x, x, x, x, 100, 100, 100, 65);
    bad |= tdx_last_status() != TDX_ERR_UNSUPPORTED;
    cp_als4(x, x, x, x, x, 10, 10, 10, 10, 65);
    bad |= (tdx_last_status() != TDX_ERR_UNSUPPORTED) << 1;
    cp_tensorcore(x, x, x, x, 100, 100, 100, 65);
    bad |= (tdx_last_status() != TDX_ERR_UNSUPPORTED) << 2;
    tucker_hosvd(x, x, x, x, x, 129, 64, 64, 8, 8, 8);
    bad |= (tdx_last_status() != TDX_ERR_UNSUPPORTED) << 3;
    tucker_tensorcore(x, x, x, x, x, 64, 129, 64, 8, 8, 8);
    bad |= (tdx_last_status() != TDX_ERR_UNSUPPORTED) << 4;
    tucker_hosvd4(x, x, x, x, x, x, 16, 16, 16, 129, 4, 4, 4, 4);
    bad |= (tdx_last_status() != TDX_ERR_UNSUPPORTED) << 5;
    return bad;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_limits_rejected_and_printed(tmp_path):
    """rank 65 (CP) and a 129-long mode (Tucker) return TDX_ERR_UNSUPPORTED through the void
    wrappers, which print the status to stderr and leave the outputs unwritten."""
    src = tmp_path / "limits.cpp"
    exe = tmp_path / "limits"
    src.write_text(LIMITS)
    libdir = os.path.dirname(tdx.LIB_PATH)
    cmd = ["g++", "-std=c++11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           str(src), "-o", str(exe), "-L", libdir, "-ltdx", "-Wl,-rpath," + libdir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stderr)
    lines = [ln for ln in r.stderr.splitlines() if ln.startswith("cudecompose: ")]
    names = [ln.split(":")[1].strip() for ln in lines]
    assert names == ["cp_als", "cp_als4", "cp_tensorcore", "tucker_hosvd", "tucker_tensorcore", "tucker_hosvd4"], r.stderr
    assert all("status %d" % tdx.TDX_ERR_UNSUPPORTED in ln for ln in lines)
