"""Writes the translation units that instantiate the kernel variants.

Run by build.py before compiling; each generated file registers the
launchers it contains in the kernel table (registry.hpp).
"""
import itertools
import os

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "generated")

# operator variant: (dim, n, nc, fp32, geo, skew, diff, fin, epi); geo 0
# reads the precomputed geometry tables, geo 1 evaluates axis-aligned
# geometry and the speed per point (SIPG with a variable speed), geo 2 is
# axis-aligned with a constant speed (products of 1D weights)
FIN_INVMASS, FIN_INTEGRATE, FIN_LOCAL = 0, 1, 2
EPI_STORE, EPI_LSRK = 0, 1
GEO_TABLE, GEO_ANALYTIC, GEO_UNIFORM = 0, 1, 2
ADVECTION_GEO = (GEO_TABLE, GEO_UNIFORM)
SIPG_GEO = (GEO_ANALYTIC, GEO_UNIFORM)


def op_variants():
    v = []
    # explicit right-hand side M^-1(-L u): apply and fused LSRK stage, the
    # conservative (skew 0) and the skew-symmetric form
    for dim, n, cart, epi, skew in itertools.product(
            (2, 3), range(2, 9), ADVECTION_GEO, (EPI_STORE, EPI_LSRK),
            (0, 1)):
        v.append((dim, n, 1, 0, cart, skew, 0, FIN_INVMASS, epi))
    # fp32 explicit right-hand side
    for dim, n, cart, skew in itertools.product((2, 3), range(2, 9),
                                                ADVECTION_GEO, (0, 1)):
        v.append((dim, n, 1, 1, cart, skew, 0, FIN_INVMASS, EPI_STORE))
    # implicit single-stage operator (M/dt + L) and IRK right-hand side,
    # skew-symmetric form, fp64 and fp32
    for dim, n, cart, fp32 in itertools.product(
            (2, 3), range(2, 9), ADVECTION_GEO, (0, 1)):
        v.append((dim, n, 1, fp32, cart, 1, 0, FIN_INTEGRATE, EPI_STORE))
    # stage-coupled IRK operator
    for dim, n, nc, fp32, cart in itertools.product(
            (2, 3), range(2, 8), (2, 3, 4), (0, 1), ADVECTION_GEO):
        v.append((dim, n, nc, fp32, cart, 1, 0, FIN_INTEGRATE, EPI_STORE))
    # cell-local stage operator of the batched cell solver (collocation
    # basis), 1..4 stages
    for dim, n, nc, fp32, cart in itertools.product(
            (2, 3), range(2, 8), (1, 2, 3, 4), (0, 1), ADVECTION_GEO):
        v.append((dim, n, nc, fp32, cart, 1, 0, FIN_LOCAL, EPI_STORE))
    # advection + SIPG diffusion, axis-aligned: explicit (apply, LSRK stage)
    # and implicit operator, fp64 and fp32
    for dim, n, fp32, geo in itertools.product((2, 3), range(2, 9), (0, 1),
                                               SIPG_GEO):
        v.append((dim, n, 1, fp32, geo, 0, 1, FIN_INVMASS, EPI_STORE))
        v.append((dim, n, 1, fp32, geo, 0, 1, FIN_INVMASS, EPI_LSRK))
        v.append((dim, n, 1, fp32, geo, 1, 1, FIN_INTEGRATE, EPI_STORE))
    return v


N_SHAPES = 3


def tile_shapes(var):
    """Tile shapes built for an operator variant: all three for the 3D
    explicit right-hand sides in conservative form (apply and fused LSRK
    stage, advection and SIPG, FP64 and FP32), which the autotuner chooses
    among; the default shape for the others."""
    dim, n, nc, fp32, geo, skew, diff, fin, epi = var
    if dim == 3 and nc == 1 and fin == FIN_INVMASS and not skew:
        return range(N_SHAPES)
    return (0,)


def shape_cells(lines, shape):
    """Cells per workgroup (shape_cells in op_kernel.hpp)."""
    if shape == 1:
        return 1 if lines >= 64 else 64 // lines
    if shape == 2:
        return 1 if lines >= 256 else 256 // lines
    return 1 if lines >= 128 else 128 // lines


def fdm_variants():
    v = []
    for dim, n, ns, prec in itertools.product((2, 3), range(2, 9), (1, 2, 3, 4),
                                              (0, 1, 2)):
        v.append(("fdm", dim, n, ns, prec))
    return v


FDM_TYPES = {0: ("double", "double", "double"), 1: ("float", "float", "float"),
             2: ("float", "double", "double")}


def burgers_variants():
    # 2D: the point-wise LDS kernel, degrees 1..4; 3D: the slab-wise
    # kernel, degrees 1..7
    v = []
    for dim, k, cart, epi in itertools.product((2, 3), range(1, 8), (0, 1),
                                               (EPI_STORE, EPI_LSRK)):
        if dim == 2 and k > 4:
            continue
        v.append(("burgers", dim, k + 1, k * (k + 1) // 2, cart, epi))
    return v


def mass_variants():
    # cellwise inverse mass matrix, fp64 and fp32
    return [("mass", dim, n, fp32) for dim, n, fp32 in
            itertools.product((2, 3), range(2, 9), (0, 1))]


def colloc_variants():
    # batched cell solver: collocation transform and the per-cell update
    v = []
    for dim, n, fp32 in itertools.product((2, 3), range(2, 8), (0, 1)):
        v.append(("colloc", dim, n, fp32))
        for ns in (1, 2, 3, 4):
            v.append(("cellgmres", dim, n, ns, fp32))
    return v


def batch_variants():
    # serving loop: batched explicit LSRK stages, 3D, degrees 1..7, every
    # tile shape
    return [("batch", 3, n, sh) for n in range(2, 9)
            for sh in range(N_SHAPES)]


def b(x):
    return "true" if x else "false"


def emit(path, variants):
    lines = ['#include "../registry.hpp"', "", "namespace dgops", "{"]
    for i, var in enumerate(variants):
        if var[0] == "batch":
            _, dim, n, sh = var
            cpb = shape_cells(n ** (dim - 1), sh)
            lines.append(
                f"  static BatchRegistrar s{i}(batch_key({dim}, {n}, {sh}), "
                f"&launch_batch<{dim}, {n}, {cpb}>);")
            continue
        if var[0] == "colloc":
            _, dim, n, fp32 = var
            T = "float" if fp32 else "double"
            lines.append(
                f"  static CollocRegistrar c{i}(colloc_key({dim}, {n}, {fp32}), "
                f"&launch_to_colloc<{dim}, {n}, {T}>);")
            continue
        if var[0] == "cellgmres":
            _, dim, n, ns, fp32 = var
            T = "float" if fp32 else "double"
            lines.append(
                f"  static CellGmresRegistrar q{i}(cell_gmres_key({dim}, {n}, "
                f"{ns}, {fp32}), &launch_cell_gmres<{dim}, {n}, {ns}, {T}>);")
            continue
        if var[0] == "mass":
            _, dim, n, fp32 = var
            T = "float" if fp32 else "double"
            lines.append(
                f"  static MassRegistrar m{i}(mass_key({dim}, {n}, {fp32}), "
                f"&launch_mass<{dim}, {n}, {T}>);")
            continue
        if var[0] == "burgers":
            _, dim, n, m, cart, epi = var
            fn = "launch_burgers3" if dim == 3 else "launch_burgers"
            lines.append(
                f"  static BurgersRegistrar g{i}(burgers_key({dim}, {n}, {cart}, "
                f"{epi}), &{fn}<{dim}, {n}, {m}, {b(cart)}, {epi}>);")
            continue
        if var[0] == "fdm":
            _, dim, n, ns, prec = var
            tc, ti, to = FDM_TYPES[prec]
            lines.append(
                f"  static FdmRegistrar f{i}(fdm_key({dim}, {n}, {ns}, {prec}), "
                f"&launch_fdm<{dim}, {n}, {ns}, {tc}, {ti}, {to}>);")
            continue
        dim, n, nc, fp32, geo, skew, diff, fin, epi, sh = var
        T = "float" if fp32 else "double"
        cpb = shape_cells(n ** (dim - 1), sh)
        lines.append(
            f"  static OpRegistrar r{i}(op_key({dim}, {n}, {nc}, {fp32}, {geo}, "
            f"{skew}, {diff}, {fin}, {epi}, {sh}), &launch_op<{dim}, {n}, "
            f"{nc}, {T}, {geo}, {b(skew)}, {b(diff)}, {fin}, {epi}, {cpb}>);")
    lines += ["} // namespace dgops", ""]
    text = "\n".join(lines)
    if os.path.exists(path) and open(path).read() == text:
        return
    with open(path, "w") as f:
        f.write(text)


def main(n_files=24):
    os.makedirs(OUT, exist_ok=True)

    def weight(v):
        if v[0] == "fdm":
            return v[2] ** (v[1] + 1) * v[3]
        if v[0] == "burgers":
            return v[3] ** v[1] * v[2]
        if v[0] in ("batch", "mass", "colloc", "cellgmres"):
            return v[2] ** (v[1] + 1)
        return v[1] ** (v[0] + 1) * v[2] * (2 if v[6] else 1)
    ops = [v + (sh,) for v in op_variants() for sh in tile_shapes(v)]
    variants = (ops + fdm_variants() + burgers_variants() +
                batch_variants() + mass_variants() + colloc_variants())
    # spread heavy (3D, large n) variants evenly
    variants.sort(key=lambda v: -weight(v))
    groups = [[] for _ in range(n_files)]
    for i, v in enumerate(variants):
        groups[i % n_files].append(v)
    names = []
    for g, vs in enumerate(groups):
        name = os.path.join(OUT, f"op_{g:02d}.hip")
        emit(name, vs)
        names.append(name)
    # drop files of an earlier, larger split
    for f in os.listdir(OUT):
        p = os.path.join(OUT, f)
        if f.endswith(".hip") and p not in names:
            os.remove(p)
    return names


if __name__ == "__main__":
    for n in main():
        print(n)
