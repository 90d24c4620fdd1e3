"""Writes tests/golden/golden.json: known-answer vectors for the oracle.

Every expected value here comes from a closed form or a brute-force
definition evaluated with Python integers, not from the oracle's code:
  - field / extension arithmetic from the definitions of F_p and F_p[x]/(x^4-3);
  - roots of unity from 3^((p-1)/2^k), with orders checked;
  - DFT by its O(N^2) definition;
  - eq tables and multilinear evaluation by the product / sum definitions
    (variable 0 on the most significant index bit);
  - sumcheck round values h(0), h(2) by evaluating the multilinear
    extensions of both factors at X = 0, 2 and summing;
  - extrapolate_012 by evaluating the interpolating quadratic;
  - SplitMix64 and xoshiro256++ from their published reference outputs.
The reference crate ships no fixtures of its own for this path.

Usage: python tests/golden/make_golden.py
"""
import json
import os
import random

P = 0x7F000001
W = 3


def fmul(a, b):
    return a * b % P


def finv(a):
    return pow(a, P - 2, P)


def ef_mul(a, b):
    t = [0] * 7
    for i in range(4):
        for j in range(4):
            t[i + j] += a[i] * b[j]
    return [(t[k] + W * (t[k + 4] if k + 4 < 7 else 0)) % P for k in range(4)]


def ef_add(a, b):
    return [(x + y) % P for x, y in zip(a, b)]


def ef_sub(a, b):
    return [(x - y) % P for x, y in zip(a, b)]


def ef_scale(a, s):
    return [x * s % P for x in a]


ONE = [1, 0, 0, 0]
ZERO = [0, 0, 0, 0]


def ef_pow(a, e):
    r = ONE
    while e:
        if e & 1:
            r = ef_mul(r, a)
        a = ef_mul(a, a)
        e >>= 1
    return r


def ef_inv(a):
    return ef_pow(a, P ** 4 - 2)


def root(k):
    return pow(3, (P - 1) >> k, P)


def dft(x):
    n = len(x)
    w = root(n.bit_length() - 1)
    return [sum(x[t] * pow(w, t * k, P) for t in range(n)) % P for k in range(n)]


def eq_point(z, xbits):
    acc = ONE
    for zj, b in zip(z, xbits):
        acc = ef_mul(acc, zj if b else ef_sub(ONE, zj))
    return acc


def bits(x, n):
    return [(x >> (n - 1 - j)) & 1 for j in range(n)]


def eq_table(z, scale):
    n = len(z)
    return [ef_mul(scale, eq_point(z, bits(x, n))) for x in range(1 << n)]


def mle(f, z):
    n = len(z)
    acc = ZERO
    for x in range(1 << n):
        acc = ef_add(acc, ef_mul(f[x], eq_point(z, bits(x, n))))
    return acc


def h_at(e, w, X):
    """sum_x e(X, x) w(X, x) with e(X,.) = e0 + X (e1 - e0) (multilinear in X)."""
    h = len(e) // 2
    acc = ZERO
    for i in range(h):
        ex = ef_add(e[i], ef_scale(ef_sub(e[i + h], e[i]), X))
        wx = ef_add(w[i], ef_scale(ef_sub(w[i + h], w[i]), X))
        acc = ef_add(acc, ef_mul(ex, wx))
    return acc


def quad_through(e0, e1, e2, r):
    """value at r of the quadratic through (0,e0), (1,e1), (2,e2), solved for coefficients."""
    inv2 = finv(2)
    c0 = e0
    c2 = ef_scale(ef_add(ef_sub(e2, ef_scale(e1, 2)), e0), inv2)   # (e2 - 2e1 + e0)/2
    c1 = ef_sub(ef_sub(e1, c0), c2)
    return ef_add(ef_add(c0, ef_mul(c1, r)), ef_mul(c2, ef_mul(r, r)))


def expand(y, n):
    out = [None] * n
    cur = y
    for i in range(n - 1, -1, -1):
        out[i] = cur
        cur = ef_mul(cur, cur)
    return out


def main():
    rng = random.Random(20240607)
    rf = lambda: rng.randrange(P)
    re = lambda: [rf() for _ in range(4)]
    g = {"p": P}

    g["field_mul"] = [[a, b, fmul(a, b)] for a, b in [(rf(), rf()) for _ in range(16)] + [(P - 1, P - 1), (0, 5)]]
    g["field_inv"] = [[a, finv(a)] for a in [rf() for _ in range(8)] + [1, 2, P - 1]]
    g["ef_mul"] = [[a, b, ef_mul(a, b)] for a, b in [(re(), re()) for _ in range(16)]]
    g["ef_x4"] = ef_pow([0, 1, 0, 0], 4)   # x^4 = W
    g["ef_inv"] = [[a, ef_inv(a)] for a in [re() for _ in range(4)]]
    # two-adic generators: order exactly 2^k
    gens = {}
    for k in range(0, 25):
        w = root(k)
        assert pow(w, 1 << k, P) == 1 and (k == 0 or pow(w, 1 << (k - 1), P) != 1)
        gens[k] = w
    g["two_adic_generator"] = gens
    g["dft"] = []
    for n in (1, 2, 4, 8, 16, 32):
        x = [rf() for _ in range(n)]
        g["dft"].append([x, dft(x)])
    g["eq_table"] = []
    for n in (0, 1, 2, 3, 5):
        z = [re() for _ in range(n)]
        s = re()
        g["eq_table"].append([z, s, eq_table(z, s)])
    g["mle_base"] = []
    for n in (1, 3, 6):
        f = [rf() for _ in range(1 << n)]
        z = [re() for _ in range(n)]
        g["mle_base"].append([f, z, mle([[v, 0, 0, 0] for v in f], z)])
    g["mle_ext"] = []
    for n in (1, 4):
        f = [re() for _ in range(1 << n)]
        z = [re() for _ in range(n)]
        g["mle_ext"].append([f, z, mle(f, z)])
    g["sumcheck_coeffs"] = []
    for n in (1, 2, 5):
        e = [re() for _ in range(1 << n)]
        w = [re() for _ in range(1 << n)]
        g["sumcheck_coeffs"].append([e, w, h_at(e, w, 0), h_at(e, w, 2)])
    g["fold"] = []
    for n in (1, 3):
        a = [re() for _ in range(1 << n)]
        r = re()
        h = len(a) // 2
        g["fold"].append([a, r, [ef_add(a[i], ef_mul(r, ef_sub(a[i + h], a[i]))) for i in range(h)]])
    g["extrapolate_012"] = []
    for _ in range(6):
        e0, e1, e2, r = re(), re(), re(), re()
        g["extrapolate_012"].append([e0, e1, e2, r, quad_through(e0, e1, e2, r)])
    g["expand_from_univariate"] = []
    for n in (1, 5):
        y = re()
        g["expand_from_univariate"].append([y, n, expand(y, n)])
    # SplitMix64 with state 0 (reference outputs of the published generator) and
    # xoshiro256++ from state [1, 2, 3, 4] (reference outputs of the published
    # generator, also the rand_xoshiro known-answer test).
    g["splitmix64_seed0"] = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    g["xoshiro256pp_1234"] = [41943041, 58720359, 3588806011781223, 3591011842654386, 9228616714210784205,
                              9973669472204895162, 14011001112246962877, 12406186145184390807,
                              15849039046786891736, 10450023813501588000]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(out, "w") as fh:
        json.dump(g, fh, separators=(",", ":"))
    print(out)


if __name__ == "__main__":
    main()
