"""Test infrastructure: apply the product's flattened device plan with numpy,
mirroring the HIP kernels (destination-major gather-GEMV passes plus the
near-field kernel), so that the fused operators can be compared with the
oracle on a machine without a GPU."""
import numpy as np


def apply_plan(plan, x, n_state, find_phi=True, rank=36):
    N = x.size
    st = np.zeros(n_state, dtype=complex)
    st[:N] = x
    ops, src, D = plan["ops"], plan["src"], plan["dests"]
    near_done = False

    # an adaptive tree's near field is a store pass into the output rows
    near_passes = [p for p in np.unique(D[:, 0])
                   if np.all(D[D[:, 0] == p, 1] == 0) and np.all((D[D[:, 0] == p, 4] >= N) & (D[D[:, 0] == p, 4] < 2 * N))]

    def apply(rows):
        out = []
        for _, kind, R, Cn, dst, oo, so in rows:
            A = ops[oo:oo + R * Cn].reshape(Cn, R).T
            out.append((kind, dst, R, A @ st[src[so:so + Cn]]))
        for kind, dst, R, v in out:
            if kind == 1:
                st[dst:dst + R] += v
            else:
                st[dst:dst + R] = v

    def near():
        if plan["T"].size == 0:
            for p in near_passes:
                apply(D[D[:, 0] == p])
            return
        T, nbr, qk2 = plan["T"].reshape(9, rank, rank), plan["nbr"], plan["qk2"]
        X = x.reshape(-1, rank)
        acc = np.zeros_like(X)
        for s in range(9):
            idx = nbr[:, s]
            ok = idx >= 0
            acc[ok] += X[idx[ok]] @ T[s].T
        y = X + acc * qk2.reshape(-1, rank) if find_phi else acc
        st[N:2 * N] = y.reshape(-1)

    for p in np.unique(D[:, 0]):
        if p in near_passes:
            continue
        rows = D[D[:, 0] == p]
        if rows[0, 1] == 1 and not near_done:
            near()
            near_done = True
        apply(rows)
    if not near_done:
        near()
    return st[N:2 * N].copy()
