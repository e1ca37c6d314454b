"""CPU checks of the C ABI: the library loads, exports every entry point the
header declares, and its host-side argument checks answer without a GPU."""
import ctypes
import os
import subprocess

import pytest

from cubismnova_amd import _abi


def test_library_exports_every_declared_symbol():
    L = _abi.lib()
    declared = _abi.declared_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(declared) <= exported


def test_every_declared_symbol_has_a_ctypes_signature():
    assert set(_abi.declared_symbols()) <= set(_abi.SIGNATURES)


def test_strerror_and_host_validation():
    L = _abi.lib()
    assert L.cnv_strerror(0) == b"success"
    assert L.cnv_strerror(_abi.CNV_ERR_NOT_ALLOCATED).startswith(b"FieldLab")
    d = _abi.cnv_grid_desc()
    d.ncomp, d.dtype = 1, _abi.CNV_F64
    for k in range(3):
        d.nblocks[k], d.block_cells[k], d.begin[k], d.end[k] = 0, 8, 0.0, 1.0
    h = ctypes.c_void_p()
    rc = L.cnv_grid_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == _abi.CNV_ERR_INVALID and not h.value
    assert b"nblocks" in L.cnv_last_error()
    for k in range(3):
        d.nblocks[k] = 2
    d.dtype = 7
    assert L.cnv_grid_create(ctypes.byref(d), ctypes.byref(h)) == _abi.CNV_ERR_INVALID
    assert L.cnv_grid_create(None, ctypes.byref(h)) == _abi.CNV_ERR_INVALID
    assert L.cnv_grid_destroy(None) == 0
    assert L.cnv_laplacian(None, 0, None, 0, None) == _abi.CNV_ERR_INVALID


def test_stencil_checks_mirror_reference():
    import cubismnova_amd as cnv
    with pytest.raises(RuntimeError, match="Stencil: begin_ must be <= 0"):
        cnv.Stencil(1, 2)
    with pytest.raises(RuntimeError, match="Stencil: end_ must be > 0"):
        cnv.Stencil(-1, 0)
    s = cnv.Stencil(-3, 4, True)
    assert s.getBegin() == [-3] * 3 and s.getEnd() == [4] * 3 and s.isTensorial()
    assert cnv.Absorbing(0, 1).name() == "Zeroth-Order Absorbing"
    assert cnv.Dirichlet(0, 0, 3.0).getValue() == 3.0
    assert cnv.Symmetry(2, 1).getBoundaryInfo().is_periodic is False
    assert cnv.Base(1, 0).getBoundaryInfo().is_periodic is True


def test_layout_roundtrip():
    import numpy as np
    from cubismnova_amd.layout import blocks_to_global, global_to_blocks
    u = np.arange(6 * 8 * 10, dtype=np.float64).reshape(6, 8, 10)
    nb, n = (2, 2, 3), (5, 4, 2)
    s = global_to_blocks(u, nb, n, block_stride=44)
    assert s.size == 12 * 44
    assert s[0:5].tolist() == [0, 1, 2, 3, 4]
    assert s[44] == 5  # block (1,0,0) starts at x = 5
    assert np.array_equal(blocks_to_global(s, nb, n, 44), u)


def test_tile_tables_host_side():
    """The tile tables are host state: valid shapes are accepted, shapes
    without a kernel are rejected, and nothing touches a GPU."""
    import numpy as np
    import cubismnova_amd as cnv
    base = cnv.tune_diffusion_run(np.float32, 32)
    w64 = cnv.tune_weno5(np.float64, 64)
    j64, j32 = cnv.tune_jacobi(np.float64, 64), cnv.tune_jacobi(np.float32, 64)
    try:
        cnv.set_tune_diffusion_run(np.float32, 32, 256, 8)
        assert cnv.tune_diffusion_run(np.float32, 32) == (256, 8)
        cnv.set_tune_diffusion_run(np.float32, 32, 512, 16)   # 8-byte vectors, 512 threads
        assert cnv.tune_diffusion_run(np.float32, 32) == (512, 16)
        cnv.set_tune_diffusion_run(np.float32, 16, 128, 8)    # 8-byte vectors on 16^3 blocks
        with pytest.raises(cnv.CnvError):
            cnv.set_tune_diffusion_run(np.float32, 32, 1024, 16)
        with pytest.raises(cnv.CnvError):
            cnv.set_tune_weno5(np.float32, 32, 512, 16)        # no 512-thread FP32 WENO5 tile
        cnv.set_tune_weno5(np.float64, 32, 512, 16)
        # shapes of more than eight rows per thread (they would spill
        # registers to scratch memory) have no kernel
        for dt, wg in ((np.float32, 64), (np.float64, 64), (np.float64, 128)):
            with pytest.raises(cnv.CnvError):
                cnv.set_tune_weno5(dt, 64, wg, 16)
        cnv.set_tune_weno5(np.float64, 64, 256, 16)
        with pytest.raises(cnv.CnvError):
            cnv.set_tune_jacobi(np.float64, 64, 128, 16, 0)
        cnv.set_tune_jacobi(np.float64, 64, 256, 16, 1)
        assert cnv.tune_jacobi(np.float64, 64) == (256, 16, 1)
        cnv.set_tune_jacobi(np.float32, 64, 128, 8, 0)
        assert cnv.tune_jacobi(np.float32, 64) == (128, 8, 0)
    finally:
        cnv.set_tune_diffusion_run(np.float32, 32, *base)
        cnv.set_tune_diffusion_run(np.float32, 16, 64, 16)
        cnv.set_tune_weno5(np.float64, 32, 256, 32)
        cnv.set_tune_weno5(np.float64, 64, *w64)
        cnv.set_tune_jacobi(np.float64, 64, *j64)
        cnv.set_tune_jacobi(np.float32, 64, *j32)


def test_build_rejects_kernels_with_scratch_memory():
    """build.py reads the compiler's kernel-resource-usage remarks and fails
    on any kernel with per-lane scratch memory (register spills)."""
    from cubismnova_amd import build
    remarks = "\n".join([
        "x.hip:45:1: remark: Function Name: _Z1aPf [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:45:1: remark:     VGPRs: 256 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:45:1: remark:     ScratchSize [bytes/lane]: 132 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:60:1: remark: Function Name: _Z1bPf [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:60:1: remark:     ScratchSize [bytes/lane]: 0 [-Rpass-analysis=kernel-resource-usage]",
    ])
    assert build.scratch_kernels(remarks) == [("_Z1aPf", 132)]
    assert build.scratch_kernels(remarks.replace(": 132 ", ": 0 ")) == []


def test_ctypes_structs_match_the_header(tmp_path):
    """Every struct of include/cubismnova_amd.h has the size and field
    offsets that the ctypes binding declares (a C program compiled with gcc
    against the header prints them)."""
    import ctypes
    import subprocess
    from cubismnova_amd import _abi
    structs = {"cnv_bc": _abi.cnv_bc, "cnv_stencil": _abi.cnv_stencil, "cnv_grid_desc": _abi.cnv_grid_desc,
               "cnv_solve_opts": _abi.cnv_solve_opts, "cnv_solve_info": _abi.cnv_solve_info,
               "cnv_batch_item": _abi.cnv_batch_item}
    # ctypes names that differ from the C member names (Python keywords)
    cname = {("cnv_batch_item", "grid_in"): "in", ("cnv_batch_item", "grid_out"): "out"}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cubismnova_amd.h"', 'int main(void) {']
    for name, cls in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (name, name))
        for f in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));'
                         % (name, f[0], name, cname.get((name, f[0]), f[0])))
    lines += ['return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    r = subprocess.run(["gcc", "-std=c99", "-I" + inc, str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    for name, cls in structs.items():
        assert int(got[name]) == ctypes.sizeof(cls), name
        for f in cls._fields_:
            assert int(got[name + "." + f[0]]) == getattr(cls, f[0]).offset, (name, f[0])


def test_two_sweep_jacobi_tiles_host_side():
    """The two-sweep Jacobi tile table: defaults of whole 64-plane blocks,
    shapes without a kernel (FP64 tiles of 1024 threads spill; a 64^3 block
    needs at least 256 threads for its ring; more rows than the block is
    wide) refused, valid shapes kept, no GPU touched."""
    import numpy as np
    import cubismnova_amd as cnv
    f32, f64 = cnv.get_tune_jacobi2(np.float32, 64), cnv.get_tune_jacobi2(np.float64, 64)
    assert f32 == (256, 64, 1) and f64 == (256, 64, 2)
    try:
        for dt, edge, wg in ((np.float64, 64, 1024), (np.float32, 64, 128), (np.float32, 16, 128),
                             (np.float32, 64, 96)):
            with pytest.raises(cnv.CnvError):
                cnv.set_tune_jacobi2(dt, edge, wg, 16)
        cnv.set_tune_jacobi2(np.float32, 64, 1024, 32)
        assert cnv.get_tune_jacobi2(np.float32, 64) == (1024, 32, 1)
        cnv.set_tune_jacobi2(np.float64, 64, 1024, 64, 2)     # half planes: no spill at 1024 threads
        assert cnv.get_tune_jacobi2(np.float64, 64) == (1024, 64, 2)
        cnv.set_tune_jacobi2(np.float64, 64, 256, 64, 4)      # quarter planes
        with pytest.raises(cnv.CnvError):
            cnv.set_tune_jacobi2(np.float64, 64, 512, 64, 8)
        with pytest.raises(cnv.CnvError):
            cnv.set_tune_jacobi2(np.float32, 64, 256, 0)
    finally:
        cnv.set_tune_jacobi2(np.float32, 64, *f32)
        cnv.set_tune_jacobi2(np.float64, 64, *f64)
