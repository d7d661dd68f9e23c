"""Oracle pin: mpmath brute force of alpha^T K_L beta for tiny N.

Test infrastructure only.  Sums every (i, j) term with 50-digit arithmetic,
using the reference's kernel definition (dwba_matrix_elements.py:1057-1080):
off-diagonal r_<^L / r_>^(L+1), diagonal (1-1e-12)^L / r_i.
"""
from __future__ import annotations

import mpmath as mp


def il_brute(r, alpha, beta, L: int, dps: int = 50) -> float:
    with mp.workdps(dps):
        n = len(r)
        rr = [mp.mpf(float(x)) for x in r]
        a = [mp.mpf(float(x)) for x in alpha]
        b = [mp.mpf(float(x)) for x in beta]
        diag = mp.mpf(1) - mp.mpf("1e-12") if L > 0 else mp.mpf(1)
        tot = mp.mpf(0)
        for i in range(n):
            if a[i] == 0:
                continue
            for j in range(n):
                if i == j:
                    k = diag ** L / rr[i]
                else:
                    lo, hi = (rr[i], rr[j]) if rr[i] < rr[j] else (rr[j], rr[i])
                    k = lo ** L / hi ** (L + 1)
                tot += a[i] * k * b[j]
        return float(tot)
