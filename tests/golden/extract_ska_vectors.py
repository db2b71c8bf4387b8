"""Copy the reference's PCoA / centring / PERMANOVA test vectors into ska_vectors.json.

Reads the numeric arrays of src/test_ska.cpp (test_center_mat, test_pcoa,
test_permanova_ties / _noties / _unequal) and the data file
src/test_ska_pcoa_big.dat (57 x 57 matrix, 7-dimension PCoA) of the reference
checkout given as argv[1] (default /root/reference).  Only numbers are copied.
Run once; the JSON is committed.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def body_of(src, fn):
    start = src.index(f"void {fn}()")
    return src[start:src.index("SUITE_END", start)]


def array(body, name):
    m = re.search(r"\b" + re.escape(name) + r"\[\]\s*=\s*\{(.*?)\};", body, flags=re.S)
    text = re.sub(r"//[^\n]*", "", m.group(1))
    return [float(x) for x in re.findall(r"-?\d+(?:\.\d+)?(?:[eE][-+]?\d+)?", text)]


def scalar(body, name):
    m = re.search(r"\b" + re.escape(name) + r"\s*=\s*(-?\d+(?:\.\d+)?)\s*;", body)
    return float(m.group(1))


def main():
    src = open(os.path.join(REF, "src", "test_ska.cpp")).read()
    out = {}
    b = body_of(src, "test_center_mat")
    out["center_mat"] = {"matrix": array(b, "matrix"), "n": 6, "expected": array(b, "exp"),
                         "cite": "src/test_ska.cpp:7-60"}
    b = body_of(src, "test_pcoa")
    out["pcoa"] = {"matrix": array(b, "matrix"), "n": 9, "centered": array(b, "exp2"),
                   "eigenvalues": array(b, "exp3a"), "eigenvectors": array(b, "exp3b"),
                   "samples": array(b, "exp4b"), "proportion": array(b, "exp4c"), "dims": 5,
                   "cite": "src/test_ska.cpp:62-245"}
    for key in ("ties", "noties", "unequal"):
        b = body_of(src, f"test_permanova_{key}")
        out[f"permanova_{key}"] = {
            "matrix_fp64": array(b, "matrix_fp64"), "matrix_fp32": array(b, "matrix_fp32"),
            "grouping": [int(x) for x in array(b, "grouping_equal")],
            "grouping_swapped": [int(x) for x in array(b, "grouping_equal_swapped")],
            "stat": scalar(b, "exp_stat"), "pvalue": scalar(b, "exp_pvalue"),
            "cite": f"src/test_ska.cpp test_permanova_{key}"}
    lines = open(os.path.join(REF, "src", "test_ska_pcoa_big.dat")).read().split("\n")
    vals, i = [], 0
    n = int(lines[0].split()[-1])
    nums = [float(x) for x in lines[1:1 + n * n]]
    dims = int(lines[1 + n * n].split()[-1])
    rest = [float(x) for x in lines[2 + n * n:] if x.strip()]
    out["pcoa_big"] = {"n": n, "dims": dims, "matrix": nums, "eigenvalues": rest[:dims],
                       "samples": rest[dims:dims + n * dims], "proportion": rest[dims + n * dims:dims + n * dims + dims],
                       "cite": "src/test_ska_pcoa_big.dat, src/test_ska.cpp:247-420"}
    json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ska_vectors.json"), "w"))


if __name__ == "__main__":
    main()
