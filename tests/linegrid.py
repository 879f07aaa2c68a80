"""Dense numpy statements of the line block ILU and the two-grid transfers, used
to check the HIP kernels (csrc/lineilu.hip, csrc/mg.hip) on small grids.

* ``dilu_apply``: M^-1 r for M = (P + L) P^-1 (P + U), lines of m = nx dof
  rows taken colour by colour ((j mod 2) + 2 (k mod 2) in 3D, j mod 2 in 2D),
  P_a = D_a - sum over earlier-coloured neighbours b of A_ab P_b^-1 A_ba.
* ``restrict``, ``prolong``, ``restrict_state``: the staggered-grid transfers
  between a grid and the grid coarsened by two in every direction of the
  problem's dimension (velocities on the east / north / top faces, scalars at
  cell centres; index d + dof (i + nx (j + ny k))).
"""
import numpy


def colour(line, ny, nz):
    cj, ck = (ny - 1 - line % ny) % 2, (nz - 1 - line // ny) % 2
    return 3 - (cj + 2 * ck) if nz > 1 else 1 - cj


def dilu_apply(A, r, nx, ny, nz, dof):
    A = A.toarray() if hasattr(A, 'toarray') else numpy.asarray(A)
    m = nx * dof
    N = ny * nz
    blk = lambda a, b: A[a * m:(a + 1) * m, b * m:(b + 1) * m]
    order = sorted(range(N), key=lambda a: (colour(a, ny, nz), a))
    col = [colour(a, ny, nz) for a in range(N)]
    Pinv = [None] * N
    for a in order:
        P = blk(a, a).copy()
        for b in range(N):
            if col[b] < col[a] and numpy.any(blk(a, b)):
                P -= blk(a, b) @ Pinv[b] @ blk(b, a)
        Pinv[a] = numpy.linalg.inv(P)
    y = numpy.zeros(N * m)
    for a in order:
        t = r[a * m:(a + 1) * m].copy()
        for b in range(N):
            if col[b] < col[a]:
                t -= blk(a, b) @ y[b * m:(b + 1) * m]
        y[a * m:(a + 1) * m] = Pinv[a] @ t
    z = y.copy()
    for a in reversed(order):
        t = numpy.zeros(m)
        for c in range(N):
            if col[c] > col[a]:
                t += blk(a, c) @ z[c * m:(c + 1) * m]
        z[a * m:(a + 1) * m] = y[a * m:(a + 1) * m] - Pinv[a] @ t
    return z


def _grid(v, n, dof):
    return numpy.asarray(v).reshape(n[2], n[1], n[0], dof)


def _coarse(n, dim):
    return tuple(n[a] // 2 if a < dim else n[a] for a in range(3))


def _fine_range(C, a, n, dim):
    return (2 * C, 2 * C + 1) if a < dim else (C,)


def restrict(r, n, dim, dof, gauge):
    N = _coarse(n, dim)
    R = _grid(r, n, dof)
    out = numpy.zeros((N[2], N[1], N[0], dof))
    for K in range(N[2]):
        for J in range(N[1]):
            for I in range(N[0]):
                C = (I, J, K)
                for d in range(dof):
                    s = 0.0
                    if d >= dim:
                        for k in _fine_range(K, 2, n, dim):
                            for j in _fine_range(J, 1, n, dim):
                                for i in _fine_range(I, 0, n, dim):
                                    s += R[k, j, i, d]
                    elif C[d] < N[d] - 1:
                        t = [a for a in range(3) if a != d]
                        for u in _fine_range(C[t[1]], t[1], n, dim):
                            for w in _fine_range(C[t[0]], t[0], n, dim):
                                for q, wt in ((0, 0.5), (1, 1.0), (2, 0.5)):
                                    F = [0, 0, 0]
                                    F[d], F[t[0]], F[t[1]] = 2 * C[d] + q, w, u
                                    if F[d] < n[d]:
                                        s += wt * R[F[2], F[1], F[0], d]
                    out[K, J, I, d] = s
    out = out.reshape(-1)
    if gauge >= 0:
        out[gauge] = numpy.asarray(r)[gauge]
    return out


def prolong(e, n, dim, dof):
    N = _coarse(n, dim)
    E = _grid(e, N, dof)
    out = numpy.zeros((n[2], n[1], n[0], dof))

    def val(I3, d):
        I3 = list(I3)
        sign = 1.0
        for a in range(3):
            if a == d:
                continue
            if I3[a] < 0:
                I3[a], sign = 0, -sign
            elif I3[a] >= N[a]:
                I3[a], sign = N[a] - 1, -sign
        return sign * E[I3[2], I3[1], I3[0], d]

    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                F = (i, j, k)
                C = tuple(F[a] // 2 if a < dim else F[a] for a in range(3))
                for d in range(dof):
                    if d >= dim:
                        out[k, j, i, d] = E[C[2], C[1], C[0], d]
                        continue
                    normal = [((F[d] - 1) // 2, 1.0)] if F[d] % 2 else [(F[d] // 2 - 1, 0.5), (F[d] // 2, 0.5)]
                    tang = []
                    for a in range(3):
                        if a == d:
                            continue
                        if a < dim:
                            tang.append((a, [(C[a], 0.75), (C[a] + (1 if F[a] % 2 else -1), 0.25)]))
                        else:
                            tang.append((a, [(C[a], 1.0)]))
                    s = 0.0
                    for In, wn in normal:
                        if In < 0:
                            continue
                        for I1, w1 in tang[0][1]:
                            for I2, w2 in tang[1][1]:
                                I3 = [0, 0, 0]
                                I3[d], I3[tang[0][0]], I3[tang[1][0]] = In, I1, I2
                                s += wn * w1 * w2 * val(I3, d)
                    out[k, j, i, d] = s
    return out.reshape(-1)


def restrict_state(x, n, dim, dof):
    N = _coarse(n, dim)
    X = _grid(x, n, dof)
    out = numpy.zeros((N[2], N[1], N[0], dof))
    for K in range(N[2]):
        for J in range(N[1]):
            for I in range(N[0]):
                C = (I, J, K)
                for d in range(dof):
                    vals = []
                    if d >= dim:
                        for k in _fine_range(K, 2, n, dim):
                            for j in _fine_range(J, 1, n, dim):
                                for i in _fine_range(I, 0, n, dim):
                                    vals.append(X[k, j, i, d])
                    else:
                        t = [a for a in range(3) if a != d]
                        for u in _fine_range(C[t[1]], t[1], n, dim):
                            for w in _fine_range(C[t[0]], t[0], n, dim):
                                F = [0, 0, 0]
                                F[d], F[t[0]], F[t[1]] = 2 * C[d] + 1, w, u
                                vals.append(X[F[2], F[1], F[0], d])
                    out[K, J, I, d] = numpy.mean(vals)
    return out.reshape(-1)
