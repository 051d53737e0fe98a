"""Derive Joe-Kuo initialisation data (degree s, polynomial a, m_1..m_s) for every
dimension of the reference's Sobol direction-number table and write it as a
compact C++ include.

The reference ships the 32 direction numbers of each of its 1,101 dimensions
(`sobol.cpp:16-3666`, the Joe-Kuo "old 1111" set, `sobol.cpp:1-14`).  The
engine regenerates those numbers at start-up from (s, a, m) with the Bratley-Fox
/ Joe-Kuo recurrence, so only the initialisation data is kept.  This script
recovers (s, a, m) for each dimension by searching the smallest recurrence that
reproduces all 32 numbers, checks the regeneration bit for bit, and writes
`sobol_init.inc`.  It reads the reference and is run once, by hand:

    python compfinance_amd/csrc/gen_sobol_init.py /root/reference/sobol.cpp
"""
import re
import sys
import os

import numpy as np

BITS = 32


def parse_table(path):
    src = open(path).read()
    arrs = re.findall(r"const unsigned (jkDir\d+)\[\] = \{([^}]*)\}", src)
    table = {}
    for name, body in arrs:
        vals = [int(x) for x in body.replace("\n", " ").split(",") if x.strip()]
        table[int(name[5:]) - 1] = vals
    return np.array([table[i] for i in range(BITS)], dtype=np.uint64)  # [bit][dim]


def regenerate(s, a, m):
    """Joe-Kuo recurrence -> 32 direction numbers (as Python ints)."""
    v = [0] * BITS
    for k in range(BITS):
        if k < s:
            v[k] = m[k] << (BITS - 1 - k)
        else:
            x = v[k - s] ^ (v[k - s] >> s)
            for j in range(1, s):
                if (a >> (s - 1 - j)) & 1:
                    x ^= v[k - j]
            v[k] = x
    return v


def search(mk):
    """mk[k] = m_{k+1} for k = 0..31.  Return (s, a) of the smallest recurrence."""
    mk = np.array(mk, dtype=np.int64)
    for s in range(1, 20):
        cand = np.arange(1 << (s - 1), dtype=np.int64)
        ok = np.ones(cand.shape, dtype=bool)
        for k in range(s, BITS):
            # m_k = 2^s m_{k-s} ^ m_{k-s} ^ XOR_j a_j 2^j m_{k-j}
            x = np.full(cand.shape, (mk[k - s] << s) ^ mk[k - s], dtype=np.int64)
            for j in range(1, s):
                bit = (cand >> (s - 1 - j)) & 1
                x ^= bit * (mk[k - j] << j)
            ok &= x == mk[k]
            if not ok.any():
                break
        if ok.any():
            return s, int(cand[np.argmax(ok)])
    raise RuntimeError("no recurrence found")


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/sobol.cpp"
    tab = parse_table(path)
    ndim = tab.shape[1]
    rows = []
    for d in range(ndim):
        col = [int(tab[k, d]) for k in range(BITS)]
        mk = [col[k] >> (BITS - 1 - k) for k in range(BITS)]
        # every direction number must be m_k shifted into the top k+1 bits
        assert all((mk[k] << (BITS - 1 - k)) == col[k] for k in range(BITS)), d
        if d == 0:
            s, a = 0, 0  # first dimension: van der Corput, all m_k = 1, marked by degree 0
            assert all(m == 1 for m in mk)
            m_init = []
            regen = [1 << (BITS - 1 - k) for k in range(BITS)]
        else:
            s, a = search(mk)
            m_init = mk[:s]
            regen = regenerate(s, a, m_init)
        assert regen == col, f"dimension {d} does not regenerate"
        rows.append((s, a, m_init))
    maxs = max(r[0] for r in rows)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sobol_init.inc")
    with open(out, "w") as f:
        f.write("// Joe-Kuo initialisation data {s, a, m_1..m_s} for %d dimensions,\n" % ndim)
        f.write("// recovered from the reference's direction-number table by gen_sobol_init.py.\n")
        f.write("// Dimension 0 (s = 0) is the van der Corput dimension: v_k = 2^(31 - k).\n")
        f.write("#define CF_SOBOL_MAX_DIM %d\n#define CF_SOBOL_MAX_DEG %d\n" % (ndim, maxs))
        f.write("static const unsigned kSobolInit[CF_SOBOL_MAX_DIM][2 + CF_SOBOL_MAX_DEG] = {\n")
        for s, a, m in rows:
            vals = [s, a] + list(m) + [0] * (maxs - len(m))
            f.write("  {" + ",".join(str(v) for v in vals) + "},\n")
        f.write("};\n")
    print("wrote", out, "dims", ndim, "max degree", maxs)


if __name__ == "__main__":
    main()
