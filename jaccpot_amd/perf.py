"""Algorithmic work counts used for roofline reporting.

M2L: F(p) = 4 S(p) + 2 Z(p) flops per pair with S(p) = sum_{l<=p} (2l+1)^2 and
Z(p) = T(p+1) + 2 sum_{m=1}^{floor(p/2)} T(p-2m+1), T(k) = k(k+1)/2:
F(2)=156, F(4)=718, F(6)=1964, F(8)=4166.
P2P: 20 flops per pairwise interaction (acceleration only).
MLP: 2 (8*64 + 64*64 + 64*3) = 9600 flops per pair forward.
"""

from __future__ import annotations


def m2l_flops(p: int) -> int:
    S = (p + 1) * (2 * p + 1) * (2 * p + 3) // 3

    def T(k):
        return k * (k + 1) // 2

    return 4 * S + 2 * (T(p + 1) + 2 * sum(T(p - 2 * m + 1) for m in range(1, p // 2 + 1)))


P2P_FLOPS_PER_INTERACTION = 20
MLP_FORWARD_FLOPS = 2 * (8 * 64 + 64 * 64 + 64 * 3)
