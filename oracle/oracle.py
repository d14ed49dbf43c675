"""NumPy restatement of the reference's stage 1 / stage 2 -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.

Sign convention (src/gravitation/kernel/py1.py:57-64): for a pair (1, 2) the reference forms
``relative_r = r1 - r2``, normalises it and does ``a1 -= relative_r * G m2 / d^2`` and
``a2 += relative_r * G m1 / d^2``; i.e. a_i = G sum_{j != i} m_j (r_j - r_i) / |r_j - r_i|^3.
The all-pairs form with G factored out is src/gravitation/kernel/pc2.py:60-90.
"""

import math

import numpy as np


def accel_f64(pos, m, G, eps=0.0, rows=None, block=8):
    """FP64 accelerations a_i for i in ``rows`` (default: all), shape (len(rows), D) for
    D-dimensional positions (as py1.py:57-64, which zips over the components).

    a_i = G * sum_{j != i} m_j (r_j - r_i) (|r_j - r_i|^2 + eps^2)^(-3/2)
    (py1.py:57-64 for eps = 0; eps > 0 adds Plummer softening, where the j = i term is
    0 * finite = 0).  The j-sum runs along a contiguous axis, so NumPy uses pairwise
    summation.
    """
    pos = np.asarray(pos, dtype=np.float64)
    m = np.asarray(m, dtype=np.float64)
    n = pos.shape[0]
    idx = np.arange(n) if rows is None else np.asarray(rows, dtype=np.int64)
    pt = np.ascontiguousarray(pos.T)  # (3, n)
    out = np.empty((idx.shape[0], pos.shape[1]), dtype=np.float64)   # any dimension, as py1
    eps2 = float(eps) * float(eps)
    for b0 in range(0, idx.shape[0], block):
        ii = idx[b0:b0 + block]
        d = pt[None, :, :] - pos[ii][:, :, None]          # (B, 3, n): r_j - r_i
        r2 = np.einsum('bkn,bkn->bn', d, d) + eps2          # (B, n)
        if eps2 == 0.0:
            r2[np.arange(ii.shape[0]), ii] = np.inf          # excludes j == i
        w = m[None, :] / (r2 * np.sqrt(r2))                  # m_j (r2)^(-3/2); 1/inf = 0
        out[b0:b0 + ii.shape[0]] = G * (d * w[:, None, :]).sum(axis=-1)
    return out


def accel_py1(pos, m, G):
    """Pure-Python FP64 pair loop in the reference's own order (py1.py:57-69): every unordered
    pair once, Newton's third law, Python floats, any dimension D.  For tiny N only."""
    n = len(pos)
    dim = len(pos[0])
    a = [[0.0] * dim for _ in range(n)]
    r = [[float(c) for c in p] for p in pos]
    mm = [float(x) for x in m]
    for i in range(n - 1):
        for j in range(i + 1, n):
            rel = [r[i][k] - r[j][k] for k in range(dim)]
            dist = math.sqrt(sum(c * c for c in rel))
            rel = [c / dist for c in rel]
            f1 = G * mm[j] / (dist ** 2)
            f2 = G * mm[i] / (dist ** 2)
            a[i] = [a[i][k] - rel[k] * f1 for k in range(dim)]
            a[j] = [a[j][k] + rel[k] * f2 for k in range(dim)]
    return np.array(a, dtype=np.float64)


def stage2(r, v, a, T):
    """Semi-implicit Euler in the reference order (src/gravitation/kernel/_base_.py:58-62,
    np2.py:110-115): v <- a*T + v, then r <- v*T + r with the new v.  Returns (r, v).
    The arithmetic type is the result type of the operands (NumPy rules)."""
    v = a * T + v
    r = v * T + r
    return r, v


def simulate(r, v, m, G, eps, dt, steps, dtype=np.float64, accel=None):
    """Oracle trajectory: ``steps`` x (accel_f64 on positions rounded to ``dtype``, stage 2 in
    FP64).  ``accel`` replaces accel_f64 by a function of the same contract (the C oracle's,
    oracle.ref.oracle_accel_c).  Returns (r, v, per-step accelerations list)."""
    accel = accel or accel_f64
    r = np.array(r, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    accs = []
    for _ in range(steps):
        rr = r.astype(dtype).astype(np.float64)
        mm = np.asarray(m).astype(dtype).astype(np.float64)
        a = accel(rr, mm, G, eps)
        accs.append(a)
        r, v = stage2(r, v, a, dt)
    return r, v, accs


def vjp_accel_f64(pos, m, mu, G, eps=0.0, block=16):
    """Vector-Jacobian products of a(r, m) = accel_f64 with the cotangent mu (N, 3):
    returns (J_r^T mu (N, 3), J_m^T mu (N,)).

    With d_kj = r_j - r_k, q = |d|^2 + eps^2, s = q^(-3/2) and the symmetric
    T_kj = s (I - 3 d d^T / q):
      (J_r^T mu)_k = G sum_{j != k} T_kj (m_k mu_j - m_j mu_k)
      (J_m^T mu)_k = G sum_{i != k} s_ik (r_k - r_i) . mu_i
    Rows k are evaluated ``block`` at a time.
    """
    pos = np.asarray(pos, dtype=np.float64)
    m = np.asarray(m, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    n = pos.shape[0]
    jr = np.empty((n, 3))
    jm = np.empty(n)
    eps2 = float(eps) ** 2
    for k0 in range(0, n, block):
        kk = np.arange(k0, min(n, k0 + block))
        d = pos[None, :, :] - pos[kk][:, None, :]          # (B, n, 3): r_j - r_k
        q = np.einsum('bjc,bjc->bj', d, d) + eps2
        if eps2 == 0.0:
            q[np.arange(kk.shape[0]), kk] = np.inf
        s = q ** -1.5
        w = m[kk][:, None, None] * mu[None] - m[None, :, None] * mu[kk][:, None, :]
        dw = np.einsum('bjc,bjc->bj', d, w)                # (m_k mu_j - m_j mu_k) . d
        jr[kk] = G * (s[:, :, None] * (w - 3.0 * (dw / q)[:, :, None] * d)).sum(axis=1)
        jm[kk] = -G * (s * np.einsum('bjc,jc->bj', d, mu)).sum(axis=1)
    return jr, jm


def trajectory_with_grad(r0, v0, m, G, eps, dt, steps, target, dtype=np.float64,
                         target_v=None, accel=None, vjp=None):
    """Oracle trajectory and the adjoint gradient of L = sum |r_T - target|^2 (plus
    sum |v_T - target_v|^2 when ``target_v`` is given; ``target`` None drops the position
    term) with respect to (r0, v0, m).  The forward step is ``simulate``'s (accelerations on
    positions rounded to ``dtype``, FP64 stage 2); the backward recursion, with
    mu = nu + dt*rho, is
      rho <- rho + dt J_r^T(r_t) mu,  gamma <- gamma + dt J_m^T(r_t) mu,  nu <- mu
    starting from rho = dL/dr_T = 2 (r_T - target), nu = dL/dv_T = 2 (v_T - target_v) (0
    without a velocity term), gamma = 0.  ``accel`` / ``vjp`` replace accel_f64 /
    vjp_accel_f64 by functions of the same contracts (the C oracle's, oracle.ref).
    Returns (L, r_T, v_T, dL/dr0, dL/dv0, dL/dm)."""
    accel = accel or accel_f64
    vjp = vjp or vjp_accel_f64
    r = np.array(r0, dtype=np.float64)
    v = np.array(v0, dtype=np.float64)
    mm = np.asarray(m).astype(dtype).astype(np.float64)
    traj = []
    for _ in range(steps):
        rr = r.astype(dtype).astype(np.float64)
        traj.append(rr)
        a = accel(rr, mm, G, eps)
        r, v = stage2(r, v, a, dt)
    loss, rho, nu = _terminal(r, v, target, target_v)
    rho, nu, gamma = adjoint_backward(traj, mm, G, eps, dt, rho, nu, vjp)
    return loss, r, v, rho, nu, gamma


def adjoint_backward(traj, m, G, eps, dt, rho, nu, vjp=None):
    """The discrete adjoint recursion of ``trajectory_with_grad`` along a given trajectory:
    ``traj[t]`` (n, 3) are the positions step t's accelerations were computed from, m the
    masses they used, (rho, nu) = (dL/dr_T, dL/dv_T).  Per step t = T-1 .. 0, with
    mu = nu + dt*rho: rho <- rho + dt J_r^T(traj[t]) mu, gamma <- gamma + dt J_m^T mu,
    nu <- mu.  Returns (dL/dr0, dL/dv0, dL/dm)."""
    vjp = vjp or vjp_accel_f64
    rho = np.array(rho, dtype=np.float64)
    nu = np.array(nu, dtype=np.float64)
    gamma = np.zeros(rho.shape[0])
    for t in range(len(traj) - 1, -1, -1):
        mu = nu + dt * rho
        jr, jm = vjp(traj[t], m, mu, G, eps)
        rho = rho + dt * jr
        gamma = gamma + dt * jm
        nu = mu
    return rho, nu, gamma


def _terminal(r, v, target, target_v):
    """The trajectory loss and its gradients with respect to r_T and v_T."""
    loss = 0.0
    rho = np.zeros_like(r)
    nu = np.zeros_like(v)
    if target is not None:
        loss += float(((r - target) ** 2).sum())
        rho = 2.0 * (r - target)
    if target_v is not None:
        loss += float(((v - target_v) ** 2).sum())
        nu = 2.0 * (v - target_v)
    return loss, rho, nu


def trajectory_loss(r0, v0, m, G, eps, dt, steps, target, target_v=None):
    """L of ``trajectory_with_grad`` on the FP64 oracle trajectory (for finite
    differences)."""
    r, v, _ = simulate(r0, v0, m, G, eps, dt, steps, dtype=np.float64)
    return _terminal(r, v, target, target_v)[0]


def rel_err(a, ref):
    """Normwise relative error ||a - ref||_2 / ||ref||_2 over all bodies and components."""
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.linalg.norm(a - ref) / np.linalg.norm(ref))


# ------------------------------------------------------------------------------------------
# CPU baselines: restatements of the reference's NumPy kernels


def np1_rows(pos, m, G, rows, dtype=np.float32, a=None):
    """The np1 algorithm (src/gravitation/kernel/np1.py:81-108) for the given rows: row i
    against every j > i, one vectorised pass per row, Newton's third law (a_i -= sum, a_j +=
    terms), in ``dtype``.  Returns the accumulated acceleration array (N, 3)."""
    pos = np.asarray(pos, dtype=dtype)
    m = np.asarray(m, dtype=dtype)
    n = pos.shape[0]
    if a is None:
        a = np.zeros((n, 3), dtype=dtype)
    g = dtype(G)
    one = dtype(1.0)
    for i in rows:
        if i >= n - 1:
            continue
        rel = pos[i] - pos[i + 1:]                       # r_i - r_j, (k, 3)
        d2 = np.add.reduce(rel * rel, axis=1)            # |r_i - r_j|^2
        unit = rel * (one / np.sqrt(d2))[:, None]
        f = g / d2
        a[i] -= np.add.reduce(unit * (f * m[i + 1:])[:, None], axis=0)
        a[i + 1:] += unit * (f * m[i])[:, None]
    return a


def np1_accel(pos, m, G, dtype=np.float32):
    """Full np1 stage 1 (all rows)."""
    n = np.asarray(pos).shape[0]
    return np1_rows(pos, m, G, range(n - 1), dtype)


def equal_pair_batches(n, workers):
    """Row ranges [(start, end)] with about equal pair counts sum(N-1-i) per batch, the split
    of src/gravitation/kernel/np4.py:103-118, made safe for small N (the reference asserts
    :115 and fails when N is small relative to the worker count)."""
    workers = max(1, min(int(workers), max(1, n - 1)))
    total = n * (n - 1) // 2
    target = total / workers
    out, start, acc = [], 0, 0
    for i in range(n - 1):
        acc += n - 1 - i
        if acc >= target * (len(out) + 1) and len(out) < workers - 1:
            out.append((start, i + 1))
            start = i + 1
    out.append((start, n - 1))
    return [b for b in out if b[1] > b[0]]


def _np4_batch(pos, m, G, rows, dtype):
    return np1_rows(pos, m, G, rows, dtype)


def np4_rows(pos, m, G, rows, workers, dtype=np.float32):
    """The np4 algorithm restricted to the listed rows (ascending): they are cut into
    ``workers`` contiguous groups of about equal pair count sum(N-1-i), one joblib process
    each, and the per-worker acceleration arrays are summed on the host
    (src/gravitation/kernel/np4.py:103-118, :150-163).  With every row this is np4_accel; a
    strided subset times a sample whose cost scales with its pair count."""
    import joblib
    n = np.asarray(pos).shape[0]
    rows = np.asarray([i for i in rows if i < n - 1], dtype=np.int64)
    cost = np.cumsum(n - 1 - rows, dtype=np.float64)
    workers = max(1, min(int(workers), rows.shape[0]))
    cuts = np.searchsorted(cost, cost[-1] * np.arange(1, workers) / workers, side='right')
    groups = [g for g in np.split(rows, cuts) if g.shape[0]]
    parts = joblib.Parallel(n_jobs=len(groups), prefer='processes')(
        joblib.delayed(_np4_batch)(pos, m, G, g, dtype) for g in groups)
    out = parts[0]
    for q in parts[1:]:
        np.add(out, q, out=out)
    return out


def np4_accel(pos, m, G, workers, dtype=np.float32):
    """The np4 algorithm (src/gravitation/kernel/np4.py:59-166): equal-pair row batches on
    ``workers`` joblib processes, each with its own full copy of the state, then a host sum of
    the per-worker acceleration arrays (:162-163)."""
    import joblib
    n = np.asarray(pos).shape[0]
    batches = equal_pair_batches(n, workers)
    parts = joblib.Parallel(n_jobs=len(batches), prefer='processes')(
        joblib.delayed(_np4_batch)(pos, m, G, range(s, e), dtype) for s, e in batches)
    out = parts[0]
    for p in parts[1:]:
        np.add(out, p, out=out)
    return out
