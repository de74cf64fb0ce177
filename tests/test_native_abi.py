"""The C-ABI library loads and exports every function include/*.h declares;
contexts are created over CPU-resident tables (no kernel runs without a GPU)."""

import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(fs\w*)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from hybrid import _native as nat
    lib = nat.lib()
    declared = _declared()
    assert "fsg_create" in declared and "fsg_rk_stage" in declared
    for name in declared:
        assert hasattr(lib, name), name
    header = (ROOT / "include" / "fluid_gfx950.h").read_text()
    want = int(re.search(r"#define FSG_ABI_VERSION (\d+)", header).group(1))
    assert lib.fsg_abi_version() == want == nat.ABI_VERSION


def test_context_over_cpu_tables_and_argument_checks():
    from conftest import PLANE_NAMES
    from hybrid import HybridCompressibleSolver, HybridNavierStokesOperator, box_mesh
    from hybrid import _native as nat
    m = box_mesh(3, split_fraction=0.4, seed=3, plane_names=PLANE_NAMES)
    s = HybridCompressibleSolver(m, polynomial_degree=2, dynamic_viscosity=0.01, device="cpu")
    op = HybridNavierStokesOperator(s)
    dm = op.device_mesh
    assert dm.ctx.value
    # every interior face contributes one record to each side, boundary faces one
    f = s.mesh.faces
    nrec = sum(int(b["rec_ptr"][-1]) for b in dm.block_tables)
    assert nrec == 2 * int((f.neighbor_block >= 0).sum()) + int((f.neighbor_block < 0).sum())
    # a record holds one face's support nodes, at most (p + 1)^2: the element
    # kernels' prefetched record registers are sized for that
    assert max(int(b["rec_nf"].max()) for b in dm.block_tables) <= (2 + 1) ** 2
    # unsupported configuration is rejected with a message, not a crash
    p = nat.Params()
    p.degree, p.nvars, p.dtype = 4, 5, 0
    import ctypes
    ctx = ctypes.c_void_p()
    rc = nat.lib().fsg_create(ctypes.byref(p), 0, None, 0, None, None, None, ctypes.byref(ctx))
    assert rc != 0 and b"degree" in nat.lib().fsg_last_error()


def test_support_tables_are_face_local():
    """Each trace table's support nodes lie on the face and carry the whole
    trace: the basis of every other node vanishes at the face points."""
    from hybrid.device import _support
    from hybrid.elements import LOCAL_FACES, basis, corners, face_quadrature
    for t in ("hexahedron", "wedge", "tetra"):
        for p in (1, 2, 3):
            for lf, fv in enumerate(LOCAL_FACES[t]):
                sup = _support(t, lf, p)
                ft = "quad" if len(fv) == 4 else "triangle"
                S, _ = face_quadrature(ft, p)
                pts = S @ corners(t)[list(fv)]
                phi = basis(t, p, pts)
                off = np.setdiff1d(np.arange(phi.shape[1]), sup)
                assert np.abs(phi[:, off]).max() < 1e-13
                assert len(sup) == ((p + 1) ** 2 if ft == "quad" else (p + 1) * (p + 2) // 2)


def test_null_arguments_are_rejected_before_any_device_work():
    """Every entry point checks its pointers first and reports through its
    module's last-error string (runs without a GPU: nothing is launched)."""
    import ctypes
    from conftest import PLANE_NAMES
    from hybrid import HybridCompressibleSolver, HybridNavierStokesOperator, box_mesh
    from hybrid import _native as nat
    lib = nat.lib()
    assert lib.fsg_diag_reset(None, None) == -10
    assert b"null" in lib.fsg_last_error()
    assert lib.fsg_residual_masked(None, None, None, 0, 0, None, None, None) == -10
    assert lib.fsg_mood_positivity(None, None, None, None) == -10
    assert lib.fsg_chem_integrate(None, None, None, None, 1e-7, 12, 0, None, None, 0, None) == -10
    assert b"null" in lib.fsg_chem_last_error()
    assert lib.fsg_mlp_adam(None, None, None, None, 0, 1e-3, 0.9, 0.999, 1e-8, 1, None) == -10
    assert b"null" in lib.fsg_mlp_last_error()
    # the masked residual names the block whose mask is missing
    m = box_mesh(3, split_fraction=0.4, seed=3, plane_names=PLANE_NAMES)
    s = HybridCompressibleSolver(m, polynomial_degree=2, dynamic_viscosity=0.01, device="cpu")
    dm = HybridNavierStokesOperator(s).device_mesh
    nb = len(dm.block_tables)
    arr = (ctypes.c_void_p * nb)()
    masks = (ctypes.c_void_p * nb)()
    rc = lib.fsg_residual_masked(dm.ctx, arr, arr, 0, 3, masks, None, None)
    assert rc == -10 and b"null mask of block 0" in lib.fsg_last_error()


def test_surrogate_inference_rejects_overlapping_output():
    """fsg_mlp_infer stages the network outputs in rho_out before it reads
    rho_in again: overlapping buffers are refused before any launch."""
    import ctypes
    from hybrid import _native as nat
    lib = nat.lib()
    n = 100
    buf = (ctypes.c_double * (2 * n * 11))()
    base = ctypes.addressof(buf)
    one = ctypes.c_void_p(base)
    for off in (0, 8 * 11 * (n - 1)):      # same buffer, and a partial overlap
        rc = lib.fsg_mlp_infer(one, one, one, ctypes.c_void_p(base), one, None,
                               ctypes.c_double(1e-7), ctypes.c_void_p(base + off), one,
                               ctypes.c_int64(n), 0, None)
        assert rc == -11 and b"overlaps" in lib.fsg_mlp_last_error()


def test_xla_custom_call_opaque():
    """The custom-call target's opaque string: the operator's own string
    decodes to its context, boundary, mode, terms and block count; malformed
    or out-of-range strings are refused with a message."""
    import ctypes
    from conftest import PLANE_NAMES
    from hybrid import FarfieldWallBoundary, HybridCompressibleSolver, \
        HybridNavierStokesOperator, box_mesh
    from hybrid import _native as nat
    m = box_mesh(3, split_fraction=0.4, seed=3, plane_names=PLANE_NAMES)
    s = HybridCompressibleSolver(m, polynomial_degree=2, dynamic_viscosity=0.01, device="cpu")
    op = HybridNavierStokesOperator(s)
    b = FarfieldWallBoundary(s, np.array([1.2, 0.9, 0.1, -0.05, 1.1]))
    text = op.xla_opaque(b)
    cfg = nat.XlaRhsConfig()
    assert nat.lib().fsg_xla_parse(text, len(text), ctypes.byref(cfg)) == 0
    assert cfg.ctx == op.device_mesh.ctx.value
    assert cfg.boundary == ctypes.addressof(op._xla_boundary)
    assert (cfg.mode, cfg.terms, cfg.nblocks) == (1, 3, 3)
    for bad in (b"ctx=0;bnd=0;mode=1;terms=3;nblocks=3", b"ctx=12;bnd=0;mode=1;terms=3",
                b"ctx=12;bnd=0;mode=7;terms=3;nblocks=3", b"ctx=zz;bnd=0;mode=1;terms=3;nblocks=3"):
        assert nat.lib().fsg_xla_parse(bad, len(bad), ctypes.byref(cfg)) == -11
        assert b"opaque" in nat.lib().fsg_last_error()
