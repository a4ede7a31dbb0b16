"""CPU-side checks of the product: the C-ABI library loads and exports every
declared entry point (no compute without a GPU), the host-side grid and field
generators, and the C++ drop-in headers compile against the ABI."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(mhs_[a-z_0-9]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol(pkg):
    import ctypes
    lib = ctypes.CDLL(pkg.mhs.LIBPATH)
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(pkg.mhs.EXPORTS) | set(pkg.transport.EXPORTS) | set(pkg.deck.EXPORTS)


def test_default_options_match_reference_defaults(pkg):
    import ctypes
    o = pkg.mhs.Options()
    pkg.mhs.lib().mhs_default_options(ctypes.byref(o))
    # IncompFlowSolverHybrid::solve defaults (:658-668)
    assert o.tol == 1e-8 and o.linsolver_type == 1 and o.maxit == 0
    assert o.prolongate_factor == 1.6 and o.smooth_steps == 1 and o.same_matrix == 0


def test_cartesian_geometry_identities(pkg):
    g = pkg.grids.cartesian(4, 3, 2, 0.5, 1.5, 2.0)
    hc = g.hf_cell()
    an = g.face_normal * g.face_area[:, None]
    # closed surfaces: sum a n = 0; sum a n (x_f - x_c)' = |c| I
    s = np.zeros((g.num_cells, 3))
    np.add.at(s, hc, an)
    assert np.abs(s).max() < 1e-14
    d = g.face_centroid - g.cell_centroid[hc]
    M = np.zeros((g.num_cells, 3, 3))
    np.add.at(M, hc, an[:, :, None] * d[:, None, :])
    assert np.allclose(M, g.cell_volume[:, None, None] * np.eye(3)[None], atol=1e-14)
    tw = g.twin()
    inside = g.nb >= 0
    assert np.all(hc[tw[inside]] == g.nb[inside])
    assert np.allclose(g.face_normal[inside], -g.face_normal[tw[inside]])


def test_periodic_partners_index_map_equals_geometric_matching(pkg):
    g = pkg.grids.cartesian(5, 4, 3, 0.3, 1.0, 2.0)
    a = pkg.grids.periodic_partners_cartesian(g, (0, 1, 2))
    b = pkg.grids.match_periodic_partners(g, [True] * 6)
    assert np.array_equal(a, b)
    a = pkg.grids.periodic_partners_cartesian(g, (1,))
    b = pkg.grids.match_periodic_partners(g, [False, False, True, True, False, False])
    assert np.array_equal(a, b)


def test_concat_instances(pkg):
    gs = [pkg.grids.cartesian(2, 2, 2), pkg.grids.cartesian(3, 1, 2)]
    u, inst = pkg.grids.concat(gs)
    assert list(inst) == [0, 8, 14]
    assert u.num_half_faces == 6 * 14
    assert np.all(u.nb[48:][u.nb[48:] >= 0] >= 8)


def test_fields_deterministic_and_distributed(pkg):
    f = pkg.fields
    a = f.normal(3, 0, 200000)
    b = f.normal(3, 0, 200000)
    assert np.array_equal(a, b)
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1) < 0.01
    assert not np.array_equal(a, f.normal(4, 0, 200000))
    u = f.uniform(1, 2, 1000, offset=500)
    assert np.array_equal(u[:10], f.uniform(1, 2, 10, offset=500))
    z = f.gaussian_filtered((32, 32, 16), 4.0, seed=9)
    assert abs(z.mean()) < 1e-12 and abs(z.std() - 1) < 1e-12
    zz = z.reshape(16, 32, 32)
    c1 = np.mean(zz * np.roll(zz, 1, axis=2))
    c8 = np.mean(zz * np.roll(zz, 12, axis=2))
    assert c1 > 0.9 and c1 > c8           # correlated at short range
    K = f.perm_c2((16, 16, 8), seed=2)
    assert np.allclose(K[:, 8], 0.1 * K[:, 0])


def test_cpp_dropin_compiles(pkg):
    import importlib.util
    spec = importlib.util.spec_from_file_location("_b", os.path.join(ROOT, "opm-porsol_amd", "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    exes = b.build_examples()
    assert any(e.endswith("known_answer_test") for e in exes)


def test_ctypes_structs_match_the_c_layout(pkg, tmp_path):
    """The ctypes mirrors of the C-ABI structs have the C compiler's size and
    field offsets (a field added on one side only would shift every later one)."""
    import ctypes
    pairs = [("mhs_solve_options", pkg.mhs.Options), ("mhs_solve_report", pkg.mhs.Report),
             ("mhs_dims", pkg.mhs.Dims), ("mhs_transport_options", pkg.transport.Options),
             ("mhs_transport_report", pkg.transport.Report), ("mhs_fluid", pkg.transport.Fluid)]
    lines = ["#include <cstddef>", "#include <cstdio>", '#include "mimetic_hybrid.h"', '#include "euler_upstream.h"',
             "int main() {"]
    for cname, py in pairs:
        lines.append(f'  std::printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  std::printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    res = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    for cname, py in pairs:
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"
