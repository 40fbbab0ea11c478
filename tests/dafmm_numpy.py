"""Independent numpy restatement of the DAFMM passes (test infrastructure).

Given the skeletons (pivot node lists) and the far-field entries of the oracle,
recompute y = A x with numpy's LU solves in place of the column-pivoted QR and
dictionaries in place of the tree objects:
  upward    LFR_M2M / HFR_M2M   GMRES/G_FMM2DTree.hpp:2387-2500, 3624-3802
  M2L       LFR_M2L / HFR_M2L   :2577-2651, 3882-3942
  downward  HFR_L2L / LFR_L2L   :4021-4236, 2702-2821
  near      evaluate_NearField_Using_Precomputations_LS :2884-2960 (dense entries)
"""
import numpy as np


def parse_skeleton(a):
    out, i = {}, 0
    while i < a.size:
        l, k, c, oa, ia, n1, n2, n3, n4, n5 = (int(v) for v in a[i:i + 10])
        i += 10
        lists = []
        for n in (n1, n2, n3, n4, n5):
            lists.append(a[i:i + n].tolist())
            i += n
        out[(l, k, c)] = dict(oa=oa, ia=ia, och=lists[0], ochg=lists[1], ich=lists[2], ichg=lists[3], ucp=lists[4])
    return out


def dafmm_numpy(o, x, rank=36, find_phi=True):
    nL, lfr, ncones = o.levels()
    sk = parse_skeleton(o.skeleton())
    nb = {l: 4 ** l for l in range(nL + 1)}
    ctr = {(l, k): o.box_center(l, k)[0] for l in range(2, nL + 1) for k in range(nb[l])}
    aper = {}
    if ncones:
        aper[0] = 2 * np.pi / ncones[0]
        for j in range(1, len(ncones)):
            aper[j] = aper[j - 1] * 2.0
    cache = {}

    def K(rows, cols):
        key = (tuple(rows), tuple(cols))
        if key not in cache:
            cache[key] = np.array([[o.far_entry(i, j) for j in cols] for i in rows], dtype=complex).reshape(len(rows), len(cols))
        return cache[key]

    def cone_toward(frm, to, level):  # cone of box `frm` facing box `to`
        d = ctr[to] - ctr[frm]
        a = np.fmod(np.arctan2(d[1], d[0]) + 2 * np.pi, 2 * np.pi)
        return int(a / aper[level])

    def is_lfr(l):
        return l >= lfr or l == nL

    def owners(l):
        if is_lfr(l):
            return [(l, k, -1) for k in range(nb[l])]
        return [(l, k, c) for k in range(nb[l]) for c in range(ncones[l])]

    def children_out(l, k, c):
        """(charge nodes, values) of the children in the order the reference concatenates."""
        nodes, vals = [], []
        for ch in range(4):
            K_ = 4 * k + ch
            key = (l + 1, K_, -1) if (c < 0 or l == lfr - 1) else (l + 1, K_, c // 2)
            nodes += sk[key]["ochg"]
            vals.append(fout[key])
        return nodes, np.concatenate(vals)

    # upward
    fout = {}
    for l in range(nL, 1, -1):
        for key in owners(l):
            s = sk[key]
            if l == nL:
                src_nodes = list(range(key[1] * rank, (key[1] + 1) * rank))
                src = x[key[1] * rank:(key[1] + 1) * rank]
            else:
                src_nodes, src = children_out(*key)
            if not s["oa"]:
                fout[key] = src
            elif s["ochg"]:
                fout[key] = np.linalg.solve(K(s["och"], s["ochg"]), K(s["och"], src_nodes) @ src)
            else:
                fout[key] = np.zeros(0, complex)
    # M2L
    u = {}
    for l in range(2, nL + 1):
        for key in owners(l):
            s = sk[key]
            u[key] = np.zeros(len(s["ucp"]), complex)
            if not s["ia"] or not s["ucp"]:
                continue
            il = o.interaction_list(l, key[1], key[2])
            for lb, kb in il:
                skey = (lb, kb, -1) if is_lfr(lb) else (lb, kb, cone_toward((lb, kb), (l, key[1]), lb))
                if sk[skey]["ochg"] and fout[skey].size:
                    u[key] += K(s["ucp"], sk[skey]["ochg"]) @ fout[skey]
    # downward
    for l in range(2, nL):
        for key in owners(l):
            s = sk[key]
            targets = []
            for ch in range(4):
                K_ = 4 * key[1] + ch
                tk = (l + 1, K_, -1) if (key[2] < 0 or l == lfr - 1 or l + 1 == nL) else (l + 1, K_, key[2] // 2)
                targets.append(tk)
            if s["ia"]:
                if not s["ichg"]:
                    continue
                fin = np.linalg.solve(K(s["ich"], s["ichg"]), u[key])
                for tk in targets:
                    if sk[tk]["ucp"]:
                        u[tk] += K(sk[tk]["ucp"], s["ichg"]) @ fin
            else:
                off = 0
                for tk in targets:
                    n = len(sk[tk]["ucp"])
                    u[tk] += u[key][off:off + n]
                    off += n
    # near field: colleague tables, y_t += qk2_t * T_s x_nbr(s) (+ x_t)
    T = o.near_tables().reshape(9, rank, rank)
    qk2 = o.contrast_k2() if find_phi else np.ones(x.size)
    offs = [(-2, -2), (0, -2), (2, -2), (2, 0), (2, 2), (0, 2), (-2, 2), (-2, 0), (0, 0)]
    r = 0.5 * ctr[(nL, 1)][0] - 0.5 * ctr[(nL, 0)][0]   # leaf radius (children 0,1 are side by side)
    where = {(round(ctr[(nL, k)][0] / r), round(ctr[(nL, k)][1] / r)): k for k in range(nb[nL])}
    y = np.zeros_like(x)
    for k in range(nb[nL]):
        cx, cy = round(ctr[(nL, k)][0] / r), round(ctr[(nL, k)][1] / r)
        acc = np.zeros(rank, complex)
        for s, (dx, dy) in enumerate(offs):
            nbk = where.get((cx + dx, cy + dy))
            if nbk is not None:
                acc += T[s] @ x[nbk * rank:(nbk + 1) * rank]
        sl = slice(k * rank, (k + 1) * rank)
        y[sl] = u[(nL, k, -1)] + qk2[sl] * acc + (x[sl] if find_phi else 0)
    return y
