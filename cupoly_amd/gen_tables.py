"""Generate the prime / root-of-i data tables used by the library and the tests.

The reference ships its RNS primes and DGT roots as generated C++ data in
``src/arithmetic/coprimes.cu`` (63-bit bucket at lines 763-866, ``NTHROOT`` at
868, ``INVNTHROOT`` at 13922, ``PROOTS`` at 26976).  BFV mode copies the first
30 entries of the 63-bit bucket into ``COPRIMES_BUCKET``
(``src/cuda/cudaengine.cu:1254-1256``, ``MAX_COPRIMES`` = 30 in
``include/cuPoly/settings.h:108-110``).  Bit-exact DGT-domain results need the
same roots, so this script reads that file as text and writes:

* ``cupoly_amd/csrc/tables63.inc``  - C initialisers compiled into the library
  (and into the CPU oracle);
* ``tests/golden/tables63.json``     - the same numbers for the Python tests.

It also writes the discrete-Gaussian cumulative table used by the sampler
(sigma = 3.2, support |e| <= 19, i.e. truncated at 6 sigma), computed with
60-digit decimal arithmetic:

* ``cupoly_amd/csrc/cdt.inc`` and ``tests/golden/cdt.json``.

Run:  python cupoly_amd/gen_tables.py [/root/reference]
"""
import decimal
import json
import os
import re
import sys

NPRIMES = 30
SIZES = [16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768]


def parse(ref_root):
    path = os.path.join(ref_root, "src", "arithmetic", "coprimes.cu")
    text = open(path).read()
    m = re.search(r"COPRIMES_63_BUCKET\[\]\s*=\s*\{([^}]*)\}", text)
    bucket = [int(x) for x in re.findall(r"\d+", m.group(1))]
    primes = bucket[:NPRIMES]

    def roots(section_name, next_name):
        start = text.index(section_name + " = {")
        end = text.index(next_name, start)
        body = text[start:end]
        out = {}
        heads = list(re.finditer(r"\{\(uint64_t\)(\d+), \{\s*\n", body))
        for h, nxt in zip(heads, heads[1:] + [None]):
            p = int(h.group(1))
            chunk = body[h.end():nxt.start() if nxt else len(body)]
            entries = re.findall(
                r"\{(\d+), \(GaussianInteger\)\{\(uint64_t\)(\d+), \(uint64_t\)(\d+)\}\}",
                chunk)
            d = {int(s): (int(r), int(i)) for s, r, i in entries}
            if p in out and out[p] != d:
                raise RuntimeError("conflicting root entries for %d" % p)
            out[p] = d
        return out

    nth = roots("NTHROOT", "INVNTHROOT")
    inv = roots("INVNTHROOT", "PROOTS")
    pstart = text.index("PROOTS = {")
    proots = {int(p): int(g) for p, g in
              re.findall(r"\{\(uint64_t\)(\d+), (\d+)\}", text[pstart:])}
    return primes, nth, inv, proots


SIGMA = "3.2"
BOUND = 19


def gaussian_cdt():
    """cdt[i] = round(2^64 * P(e <= i - BOUND)), i in [0, 2*BOUND)."""
    decimal.getcontext().prec = 60
    sigma = decimal.Decimal(SIGMA)
    w = [(-decimal.Decimal(e * e) / (2 * sigma * sigma)).exp()
         for e in range(-BOUND, BOUND + 1)]
    total = sum(w)
    cdt, acc = [], decimal.Decimal(0)
    for i in range(2 * BOUND):
        acc += w[i]
        cdt.append(int((acc / total * (1 << 64)).to_integral_value(
            rounding=decimal.ROUND_HALF_EVEN)))
    probs = [str(x / total) for x in w]
    return cdt, probs


def write_cdt(repo, here):
    cdt, probs = gaussian_cdt()
    with open(os.path.join(repo, "tests", "golden", "cdt.json"), "w") as f:
        json.dump({"sigma": SIGMA, "bound": BOUND, "cdt": [str(c) for c in cdt],
                   "probs": probs}, f, indent=0)
    lines = ["/* Generated by cupoly_amd/gen_tables.py: discrete Gaussian, sigma = %s," % SIGMA,
             " * support |e| <= %d; cdt[i] = round(2^64 * P(e <= i - %d)). */" % (BOUND, BOUND),
             "#define CUPOLY_GAUSS_BOUND %d" % BOUND,
             "static const uint64_t cupoly_gauss_cdt[%d] = {" % (2 * BOUND)]
    lines += ["  %dULL," % c for c in cdt] + ["};"]
    with open(os.path.join(here, "csrc", "cdt.inc"), "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    write_cdt(os.path.dirname(here), here)
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    primes, nth, inv, proots = parse(ref_root)
    here = os.path.dirname(os.path.abspath(__file__))
    repo = os.path.dirname(here)
    table = {"primes": [str(p) for p in primes], "sizes": SIZES,
             "proot": [], "nthroot": [], "invnthroot": []}
    for p in primes:
        table["proot"].append(proots[p])
        table["nthroot"].append([[str(nth[p][s][0]), str(nth[p][s][1])] for s in SIZES])
        table["invnthroot"].append([[str(inv[p][s][0]), str(inv[p][s][1])] for s in SIZES])
    with open(os.path.join(repo, "tests", "golden", "tables63.json"), "w") as f:
        json.dump(table, f, indent=0)

    lines = ["/* Generated by cupoly_amd/gen_tables.py from the reference's prime/root",
             " * data (src/arithmetic/coprimes.cu).  Do not edit. */",
             "#define CUPOLY_TABLE_NPRIMES %d" % len(primes),
             "#define CUPOLY_TABLE_NSIZES %d" % len(SIZES),
             "static const int cupoly_table_sizes[%d] = {%s};" % (
                 len(SIZES), ", ".join(str(s) for s in SIZES)),
             "static const uint64_t cupoly_table_primes[%d] = {" % len(primes)]
    lines += ["  %dULL," % p for p in primes] + ["};"]
    lines.append("static const uint32_t cupoly_table_proot[%d] = {%s};" % (
        len(primes), ", ".join(str(proots[p]) for p in primes)))
    for name, src in (("nthroot", nth), ("invnthroot", inv)):
        lines.append("static const uint64_t cupoly_table_%s[%d][%d][2] = {" % (
            name, len(primes), len(SIZES)))
        for p in primes:
            lines.append("  {" + ", ".join("{%dULL, %dULL}" % src[p][s] for s in SIZES) + "},")
        lines.append("};")
    with open(os.path.join(here, "csrc", "tables63.inc"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote %d primes x %d sizes" % (len(primes), len(SIZES)))


if __name__ == "__main__":
    main()
