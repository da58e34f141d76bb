"""Conjugate gradient squared as published, in FP32 numpy, for pinning the
oracle's restatement of ufo-ir's split-Bregman inner solver.

The algorithm is CGS (P. Sonneveld, SIAM J. Sci. Stat. Comput. 10 (1989) 36;
R. Barrett et al., Templates for the Solution of Linear Systems, SIAM 1994,
section 2.3.7) with the safeguards of MATLAB's cgs: breakdown flag 4 on
rho = 0 / beta = 0 / rt'v = 0 / non-finite scalars, a stagnation count when
|alpha| |uh| < eps |x|, a re-check of the true residual when the recurrence
residual is small, the count is stagnant or extra steps are running, up to
`maxmsteps` extra steps (flag 3), and on failure the return of the iterate
with the smallest checked residual when its residual is no larger.

ufo-ir's port (ufo-ir-sbtv-task.c:404-585) departs from that in four ways,
each a switch here so that a test can show the oracle follows the port and
not the textbook:
  - its "dot" and norm are the strided sum of ufo-ir-basic-ops-processor.c:
    181-193 (dot="strided"; "true" is an inner product summed in FP64);
  - the re-check computes b - A r, not b - A x (:535; recheck="r");
  - maxmsteps is 5 (:457), not min(floor(n / 50), 5, n - maxit);
  - a starting residual within tolerance returns x0 with no flag (:435-440).
Every vector operation rounds to FP32 where the port's kernels do: a + m b is
one FP32 product and one FP32 sum (operation_add2), a dense operator sums
each row in FP32 in column order."""
import math

import numpy as np

F = np.float32
EPS = 2.2204e-16  # ufo-ir-sbtv-task.c:35, compared in double


def strided_dot(a, b, w, h):
    """sum over p < h of (sum over i < w h / h of a[p i] b[p i]), each part
    and the total accumulated in FP32 in order."""
    part_len = (w * h) // h
    parts = np.zeros(h, np.float32)
    p = np.arange(h)
    for i in range(part_len):
        idx = p * i
        parts = parts + a[idx] * b[idx]
    total = F(0)
    for v in parts:
        total = F(total + v)
    return total


def true_dot(a, b, w, h):
    """Every element once, products and sum in FP64 in order, rounded once."""
    s = 0.0
    for u, v in zip(a.tolist(), b.tolist()):
        s += u * v
    return F(s)


def dense_apply(M, v):
    acc = np.zeros(M.shape[0], np.float32)
    for j in range(M.shape[1]):
        acc = acc + M[:, j] * v[j]
    return acc


def cgs(M, b, x0, maxit, w, h, dot="strided", recheck="r", maxmsteps=5, tol=1e-6):
    """M is a dense matrix or a function v -> A v on flat FP32 vectors.
    Returns (x, flag, events); events counts the branches taken and names
    the exit."""
    b = np.asarray(b, np.float32).ravel()
    x = np.asarray(x0, np.float32).ravel().copy()
    d = strided_dot if dot == "strided" else true_dot

    def norm(v):
        return F(math.sqrt(float(d(v, v, w, h))))

    if callable(M):
        def A(v):
            return np.asarray(M(v), np.float32).ravel()
    else:
        Md = np.asarray(M, np.float32)

        def A(v):
            return dense_apply(Md, v)

    ev = {"early": 0, "flag4": 0, "stag": 0, "recheck": 0, "moresteps": 0, "converged": 0, "xmin": 0,
          "restored": 0, "exit": "maxit"}
    tolb = F(F(tol) * norm(b))
    r = b - A(x)
    normr = norm(r)
    normr_act = normr
    if normr <= tolb:
        ev["early"] = 1
        ev["exit"] = "early"
        return x, 0, ev
    rt = r.copy()
    xmin = x.copy()
    normrmin = normr
    rho = F(1)
    stag = moresteps = 0
    flag = 1
    q = np.zeros_like(r)
    p = np.zeros_like(r)
    u = np.zeros_like(r)
    for it in range(maxit):
        rho1 = rho
        rho = d(rt, r, w, h)
        if rho == 0 or not np.isfinite(rho):
            flag = 4
            break
        if it == 0:
            u = r.copy()
            p = u.copy()
        else:
            beta = F(rho / rho1)
            if beta == 0 or not np.isfinite(beta):
                flag = 4
                break
            u = r + beta * q
            p = u + beta * (q + beta * p)
        vh = A(p)
        rtvh = d(rt, vh, w, h)
        if rtvh == 0:
            flag = 4
            break
        alpha = F(rho / rtvh)
        if not np.isfinite(alpha):
            flag = 4
            break
        q = u + (-alpha) * vh
        uh = u + q
        if abs(float(alpha)) * float(norm(uh)) < EPS * float(norm(x)):
            stag += 1
            ev["stag"] += 1
        else:
            stag = 0
        x = x + alpha * uh
        qh = A(uh)
        r = r + (-alpha) * qh
        normr = norm(r)
        normr_act = normr
        if normr <= tolb or stag >= 3 or moresteps:
            ev["recheck"] += 1
            r = b - A(r if recheck == "r" else x)
            normr_act = norm(r)
            if normr_act <= tolb:
                flag = 0
                ev["converged"] = 1
                ev["exit"] = "converged"
                break
            if stag >= 3 and moresteps == 0:
                stag = 0
            moresteps += 1
            ev["moresteps"] = moresteps
            if moresteps >= maxmsteps:
                flag = 3
                ev["exit"] = "moresteps"
                break
        if normr_act < normrmin:
            normrmin = normr_act
            xmin = x.copy()
            ev["xmin"] += 1
        if stag >= 3:
            flag = 3
            ev["exit"] = "stagnation"
            break
    if flag != 0:
        rcomp = b - A(xmin)
        if norm(rcomp) <= normr_act:
            x = xmin.copy()
            ev["restored"] = 1
    if flag == 4:
        ev["flag4"] = 1
        ev["exit"] = "breakdown"
    return x, flag, ev
