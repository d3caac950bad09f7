"""Provenance recipe for the HLL++ interpolation tables (run once, here).

The numbers are the empirical raw-estimate / bias tables of
Heule, Nunkesser, Hall, "HyperLogLog in Practice" (EDBT 2013), as shipped
for precisions 13..18 in include/detail/hyperloglog/tuning.h of the
reference (cuCollections lineage, Apache-2.0).  They are data, not code:
this script reads the numbers out of that header and writes them as
oracle/data/hllpp_p{13..18}.npz (for the oracle) and as
ops-collections_amd/csrc/hllpp_tables.inc (for the library's host-side
finalizer).  Neither output is derived from anything else.
"""
import os
import re

import numpy as np

SRC = "/root/reference/include/detail/hyperloglog/tuning.h"
HERE = os.path.dirname(os.path.abspath(__file__))
INC = os.path.join(HERE, "..", "..", "ops-collections_amd", "csrc", "hllpp_tables.inc")


def arr(text, name):
    m = re.search(r"\b" + name + r"\s*\{([^}]*)\}", text)
    return np.array([float(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()], np.float64)


def main():
    text = open(SRC).read()
    thr = arr(text, "threshold_data")
    lines = ["// HLL++ empirical tables (Heule et al., EDBT 2013), precisions 13..18.",
             "// Generated by oracle/data/extract_hllpp.py; data only.",
             "static const double kHllThreshold[15] = {" + ", ".join(repr(float(x)) for x in thr) + "};"]
    for p in range(13, 19):
        raw = arr(text, "raw_estimate_data_p%d" % p)
        bias = arr(text, "bias_data_p%d" % p)
        assert len(raw) == len(bias), p
        np.savez(os.path.join(HERE, "hllpp_p%d.npz" % p), raw=raw, bias=bias, threshold=thr[p - 4])
        lines.append("static const double kHllRaw%d[%d] = {%s};" % (p, len(raw), ", ".join(repr(float(x)) for x in raw)))
        lines.append("static const double kHllBias%d[%d] = {%s};" % (p, len(bias), ", ".join(repr(float(x)) for x in bias)))
    open(INC, "w").write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
