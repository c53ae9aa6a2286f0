#!/usr/bin/env python3
"""Generate the multiply-with-carry multiplier table used by every photon stream.

Method (Marsaglia, "Random number generators", J. Mod. Appl. Stat. Meth. 2(1), 2003):
a lag-1 multiply-with-carry generator with base b = 2^32 and multiplier a
walks x' = (a*x + c) mod b, c' = floor((a*x + c) / b).  Its period is the order of
b modulo p = a*b - 1.  When p is a safe prime (p = 2q + 1 with q prime) and b is a
quadratic residue mod p (b = (2^16)^2 always is), the order is exactly
q = a*2^31 - 1, i.e. about 2^62.5 for a near 2^32.

This script scans a downward from 2^32 - 1 and keeps the first COUNT values for
which both p and q are prime (deterministic Miller-Rabin for n < 2^64 with the
first twelve prime bases, Jaeschke 1993).  It writes `mwc_multipliers.h`.

Usage: python gen_mwc_multipliers.py [COUNT] > mwc_multipliers.h
"""
import sys

SMALL_PRIMES = [3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71,
                73, 79, 83, 89, 97, 101, 103, 107, 109, 113, 127, 131, 137, 139, 149]
MR_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime_u64(n: int) -> bool:
    """Deterministic Miller-Rabin for n < 2^64."""
    if n < 2:
        return False
    for p in (2,) + tuple(SMALL_PRIMES):
        if n % p == 0:
            return n == p
    d = n - 1
    r = 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in MR_BASES:
        x = pow(a, d, n)
        if x == 1 or x == n - 1:
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def is_safe_multiplier(a: int) -> bool:
    p = (a << 32) - 1
    q = (a << 31) - 1
    for sp in SMALL_PRIMES:
        if p % sp == 0 or q % sp == 0:
            return False
    return is_prime_u64(q) and is_prime_u64(p)


def generate(count: int):
    out = []
    a = 0xFFFFFFFF
    while len(out) < count:
        if is_safe_multiplier(a):
            out.append(a)
        a -= 1
    return out


def main():
    count = int(sys.argv[1]) if len(sys.argv) > 1 else 4096
    table = generate(count)
    w = sys.stdout.write
    w("// Generated by gen_mwc_multipliers.py: multipliers a with a*2^32-1 and a*2^31-1 prime.\n")
    w("// Do not edit by hand; regenerate with `python gen_mwc_multipliers.py %d`.\n" % count)
    w("#pragma once\n#include <stdint.h>\n\n")
    w("#define MCML_MWC_TABLE_SIZE %du\n\n" % count)
    w("static const uint32_t mcml_mwc_multipliers[MCML_MWC_TABLE_SIZE] = {\n")
    for i in range(0, count, 6):
        w("  " + ", ".join("%uu" % v for v in table[i:i + 6]) + ",\n")
    w("};\n")


if __name__ == "__main__":
    main()
