"""Write reference-implementation input/output vectors to tests/golden/ref_vectors.npz.

The outputs come from the reference's own CPU code (oracle/_ref/libssu_ref.so,
built by oracle/Makefile from the reference sources), so this script runs only
in the development container:

    python tests/golden/make_ref_vectors.py

Each case stores the tree (parentheses arrays), the CSR table, the method and
flags, and the stripes [0, (n+1)/2) computed by the reference, in FP64
("__fp64") and in the reference's FP32 instantiation ("__fp32").
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "unifrac-binaries_amd"))

import oracle as O  # noqa: E402
import synth  # noqa: E402

CASES = [
    # name, tips, samples, density, zipf, counts, method, vaw, alpha, bypass, normalize
    ("unw", 400, 45, 0.06, 1.1, "lognormal", "unweighted", 0, 1.0, 0, 1),
    ("unw_unnorm", 400, 45, 0.06, 1.1, "lognormal", "unweighted_unnormalized", 0, 1.0, 0, 1),
    ("unw_bypass", 400, 45, 0.06, None, "lognormal", "unweighted", 0, 1.0, 1, 1),
    ("wn", 400, 45, 0.06, None, "negbin", "weighted_normalized", 0, 1.0, 0, 1),
    ("wn_raw", 400, 45, 0.06, None, "negbin", "weighted_normalized", 0, 1.0, 0, 0),
    ("wu", 400, 45, 0.06, None, "lognormal", "weighted_unnormalized", 0, 1.0, 0, 1),
    ("gen05", 400, 45, 0.06, None, "negbin", "generalized", 0, 0.5, 0, 1),
    ("vaw_unw", 300, 33, 0.08, None, "lognormal", "unweighted", 1, 1.0, 0, 1),
    ("vaw_wn", 300, 33, 0.08, None, "negbin", "weighted_normalized", 1, 1.0, 0, 1),
    ("vaw_gen05", 300, 33, 0.08, None, "negbin", "generalized", 1, 0.5, 0, 1),
]


def main():
    out = {}
    for i, (name, tips, n, dens, zipf, counts, method, vaw, alpha, bypass, norm) in enumerate(CASES):
        seed = 7000 + 10 * i
        st, le, nm = synth.random_tree(tips, seed)
        tab = synth.random_table(tips, n, dens, seed + 1, zipf, counts)
        tree = O.Tree(st.astype(np.uint8), le, nm)
        table = O.Table(*tab)
        for prec in ("fp64", "fp32"):
            key = f"{name}__{prec}"
            out[key + "__structure"] = st.astype(np.uint8)
            out[key + "__lengths"] = le
            out[key + "__names"] = np.array(nm, dtype="U16")
            out[key + "__obs_ids"] = np.array(tab[0], dtype="U16")
            out[key + "__sample_ids"] = np.array(tab[1], dtype="U16")
            out[key + "__indptr"] = tab[2]
            out[key + "__indices"] = tab[3]
            out[key + "__data"] = tab[4]
            out[key + "__method"] = np.array(method)
            out[key + "__meta"] = np.array([vaw, alpha, bypass, norm], dtype=np.float64)
            out[key + "__out"] = O.ref_stripes(tree, table, method, fp32=(prec == "fp32"), variance_adjust=bool(vaw),
                                               alpha=alpha, bypass_tips=bool(bypass), normalize=bool(norm))
        # Faith PD from the reference's set_proportions + src/unifrac.cpp loop
        out[f"{name}__faith"] = O.ref_faith_pd(tree, table)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)


if __name__ == "__main__":
    main()
