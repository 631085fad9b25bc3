"""Emit the cumulative tables of the discrete noise samplers used by the oracle's
seeded key generator (sigma = 3.19, the reference's noise width,
src/random.rs:112-128).

ORACLE / TEST INFRASTRUCTURE ONLY.

Two tables, each T[k] = floor(2^64 * P(|X| <= k)), k = 0..41, X ~ N(0, 3.19^2):
  * ROUND: |X| rounded half away from zero (the reference maps a Normal draw
    with f64::round, src/backend/mod.rs:60-69);
  * TRUNC: |X| truncated toward zero (the reference's i32::from_f64 for the
    Gaussian LWE secret, src/random.rs:145-160).
The draws themselves are this project's own (ChaCha8 words against the table),
not rand_distr's ziggurat, so noise streams are not bit-identical to Rust's.
"""
import math
import sys

SIGMA = 3.19
KMAX = 41


def phi(x):
    return 0.5 * (1.0 + math.erf(x / (SIGMA * math.sqrt(2.0))))


def table(off):
    out = []
    for k in range(KMAX + 1):
        p = phi(k + off) - phi(-(k + off))
        out.append(min(int(p * 2.0 ** 64), 2 ** 64 - 1))
    out[-1] = 2 ** 64 - 1
    return out


def emit(prefix):
    lines = []
    for name, off in (("ROUND", 0.5), ("TRUNC", 1.0)):
        vals = table(off)
        lines.append("static const uint64_t %s_CDT_%s[%d] = {" % (prefix, name, len(vals)))
        for i in range(0, len(vals), 3):
            lines.append("    " + ", ".join("0x%016xULL" % v for v in vals[i:i + 3]) + ",")
        lines.append("};")
    return "\n".join(lines)


if __name__ == "__main__":
    print(emit(sys.argv[1] if len(sys.argv) > 1 else "PZO"))
