"""Copy the reference's stripe-layout golden data into stripe_layouts.json.

Reads the numeric arrays (stripes s1..s5 and the expected matrix or condensed
form) of the six layout tests in src/test_su.cpp:887-1316 of the reference
checkout given as argv[1] (default /root/reference).  Only numbers are
copied.  Run once; the JSON is committed.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
TESTS = ["test_unifrac_stripes_to_condensed_form_even", "test_unifrac_stripes_to_condensed_form_odd",
         "test_unifrac_stripes_to_condensed_form_odd2", "test_unifrac_stripes_to_matrix_even",
         "test_unifrac_stripes_to_matrix_odd", "test_unifrac_stripes_to_matrix_odd2"]


def numbers(body):
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return [float(x) for x in re.findall(r"-?\d+(?:\.\d+)?", body)]


def main():
    src = open(os.path.join(REF, "src", "test_su.cpp")).read()
    out = {}
    for name in TESTS:
        start = src.index(f"void {name}()")
        end = src.index("SUITE_END", start)
        body = src[start:end]
        stripes = [numbers(m) for m in re.findall(r"double s\d\[\]\s*=\s*\{(.*?)\};", body, flags=re.S)]
        exp = numbers(re.search(r"(?:double|float) exp\[\d+\]\s*=\s*\{(.*?)\};", body, flags=re.S).group(1))
        key = name.replace("test_unifrac_", "")
        out[key] = {"n_samples": len(stripes[0]), "n_stripes": len(stripes), "stripes": stripes, "expected": exp}
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "stripe_layouts.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
