"""CPU execution of a planned ghost-fill program (tests only).

Executes the operations returned by amrgpu_plan_fill with numpy, with the
semantics the device kernel implements (copy / time blend, limited linear
coarse-fine interpolation, boundary conditions, periodic copy-back), so that
the planner can be checked against the oracle's bound on the CPU."""
import numpy as np

T_EPS = 1.0e-4


def time_weights(H, t, l):
    L = H.levels[l - 1]
    alphac = (L["time"] - t) / L["possk"]
    alphai = 1.0 - alphac
    assert -T_EPS <= alphai <= 1.0 + T_EPS
    if abs(alphai - 1.0) < T_EPS:
        return "new", alphai, alphac
    assert l != H.nlev
    return ("old" if abs(alphai) < T_EPS else "blend"), alphai, alphac


def execute(H, level, t, ops, bufs, masks, mthbc):
    """Apply the program to H in place (the grids of `level`)."""
    nv = H.nvar
    arrays = []
    for b in bufs:
        if b.kind == 0:
            arrays.append(H.grid(b.level, b.patch, 0))        # (ncol, nrow, nv) view
        else:
            arrays.append(np.zeros((b.ncol, b.nrow, nv)))
    def inside(a, j0, i0, nj, ni):
        assert 0 <= j0 and 0 <= i0 and j0 + nj <= a.shape[0] and i0 + ni <= a.shape[1], \
            "access outside a buffer"

    for op in sorted(ops, key=lambda o: o.stage):
        dst = arrays[op.dst]
        if op.type in (0, 1, 3):
            inside(dst, op.dj0, op.di0, op.nj, op.ni)
        if op.type == 0:
            inside(H.grid(op.src_level, op.src, 0), op.sj0, op.si0, op.nj, op.ni)
        if op.type == 3:
            inside(arrays[op.src], 0, 0, op.nj, op.ni)
        if op.type == 0:
            mode, ai, ac = time_weights(H, t, op.src_level)
            qn = H.grid(op.src_level, op.src, 0)[op.sj0:op.sj0 + op.nj, op.si0:op.si0 + op.ni]
            qo = H.grid(op.src_level, op.src, 1)[op.sj0:op.sj0 + op.nj, op.si0:op.si0 + op.ni]
            v = qn if mode == "new" else qo if mode == "old" else qn * ai + qo * ac
            dst[op.dj0:op.dj0 + op.nj, op.di0:op.di0 + op.ni] = v
        elif op.type == 1:
            crs = arrays[op.src]
            Lf, Lc = H.levels[op.level - 1], H.levels[op.level - 2]
            dxf, dyf, dxc, dyc = Lf["hx"], Lf["hy"], Lc["hx"], Lc["hy"]
            xlower, ylower = H.domain[0], H.domain[2]
            xlow_c = xlower + op.iplo * dxc
            ylow_c = ylower + op.jplo * dyc
            m = np.frombuffer(masks, dtype=np.uint8, count=op.ni * op.nj,
                              offset=op.mask_off).reshape(op.nj, op.ni)
            for jf in range(1, op.nj + 1):
                jc = int(np.floor((jf + op.jlo - 1) / float(op.raty))) - op.jplo + 1
                eta2 = ((ylower + ((jf - 1 + op.jlo) + .5) * dyf) - (ylow_c + (jc - .5) * dyc)) / dyc
                for if_ in range(1, op.ni + 1):
                    if m[jf - 1, if_ - 1]:
                        continue
                    ic = int(np.floor((if_ + op.ilo - 1) / float(op.ratx))) - op.iplo + 1
                    eta1 = ((xlower + ((if_ - 1 + op.ilo) + .5) * dxf) - (xlow_c + (ic - .5) * dxc)) / dxc
                    a, b = ic - 1, jc - 1
                    inside(crs, b - 1, a - 1, 3, 3)
                    vp10, vm10, vc = crs[b, a + 1], crs[b, a - 1], crs[b, a]
                    vp01, vm01 = crs[b + 1, a], crs[b - 1, a]
                    dupc, dumc, ducc = vp10 - vc, vc - vm10, vp10 - vm10
                    du = np.fmin(2.0 * np.fmin(np.abs(dupc), np.abs(dumc)), 0.5 * np.abs(ducc))
                    us = du * np.copysign(1.0, ducc) * np.fmax(0.0, np.copysign(1.0, dupc * dumc))
                    dvpc, dvmc, dvcc = vp01 - vc, vc - vm01, vp01 - vm01
                    dv = np.fmin(2.0 * np.fmin(np.abs(dvpc), np.abs(dvmc)), 0.5 * np.abs(dvcc))
                    vs = dv * np.copysign(1.0, dvcc) * np.fmax(0.0, np.copysign(1.0, dvpc * dvmc))
                    dst[op.dj0 + jf - 1, op.di0 + if_ - 1] = vc + eta1 * us + eta2 * vs
        elif op.type == 3:
            dst[op.dj0:op.dj0 + op.nj, op.di0:op.di0 + op.ni] = arrays[op.src][:op.nj, :op.ni]
        elif op.type == 2:
            nxl, ibeg, nyb, jbeg = op.bc
            if nxl and mthbc[0] in (1, 3):
                for i in range(nxl):
                    dst[:, i] = dst[:, nxl] if mthbc[0] == 1 else dst[:, 2 * nxl - 1 - i]
                    if mthbc[0] == 3:
                        dst[:, i, 1] = -dst[:, i, 1]
            if ibeg and mthbc[1] in (1, 3):
                for i in range(ibeg - 1, op.ni):
                    dst[:, i] = dst[:, ibeg - 2] if mthbc[1] == 1 else dst[:, 2 * ibeg - 1 - (i + 1) - 1]
                    if mthbc[1] == 3:
                        dst[:, i, 1] = -dst[:, i, 1]
            if nyb and mthbc[2] in (1, 3):
                for j in range(nyb):
                    dst[j] = dst[nyb] if mthbc[2] == 1 else dst[2 * nyb - 1 - j]
                    if mthbc[2] == 3:
                        dst[j, :, 2] = -dst[j, :, 2]
            if jbeg and mthbc[3] in (1, 3):
                for j in range(jbeg - 1, op.nj):
                    dst[j] = dst[jbeg - 2] if mthbc[3] == 1 else dst[2 * jbeg - 1 - (j + 1) - 1]
                    if mthbc[3] == 3:
                        dst[j, :, 2] = -dst[j, :, 2]
