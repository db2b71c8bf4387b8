"""Write the reference's own test data as small fixtures under tests/golden/.

Run once in the development container (reads /root/reference, needs h5py for
the .biom files, which lives in the conda interpreter):

    /opt/conda/bin/python3.9 tests/golden/extract_reference_fixtures.py

Outputs (data only -- inputs and expected outputs):
  *.biom             the BIOM data files themselves (for the file-based entry points)
  tables.json        BIOM tables as observation-major CSR:
                     test.biom (src/), crawford.biom (ci/), absquant.biom (src/testdata/)
  trees.json         newick text of src/test.tre, src/test_faith.tre,
                     src/test_faith_shear.tre, ci/crawford.tre, src/testdata/absquant.tree
  expected.json      ci/test.ssu.exp, ci/test.ssu_vaw.exp (9x9 matrices), ci/test.faith.exp,
                     src/testdata/test.absquant.wn.dm, test.absquant_not.wn.dm (3x3)
"""
import json
import os
import shutil

import h5py

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def read_biom(path):
    with h5py.File(path, "r") as f:
        def ids(p):
            return [x.decode() if isinstance(x, bytes) else str(x) for x in f[p][:]]
        return {
            "obs_ids": ids("observation/ids"),
            "sample_ids": ids("sample/ids"),
            "indptr": [int(x) for x in f["observation/matrix/indptr"][:]],
            "indices": [int(x) for x in f["observation/matrix/indices"][:]],
            "data": [float(x) for x in f["observation/matrix/data"][:]],
        }


def read_dm(path):
    with open(path) as fh:
        rows = [ln.rstrip("\n").split("\t") for ln in fh if ln.strip()]
    header = [h for h in rows[0][1:] if h]
    ids, vals = [], []
    for r in rows[1:]:
        r = [x for x in r if x != ""] if len(r) > len(header) + 1 else r
        ids.append(r[0].strip())
        vals.append([float(x) for x in r[1:] if x.strip() != ""])
    return {"ids": ids, "matrix": vals}


def read_dm_ws(path):
    with open(path) as fh:
        rows = [ln.split() for ln in fh if ln.strip()]
    return {"ids": [r[0] for r in rows[1:]], "matrix": [[float(x) for x in r[1:]] for r in rows[1:]]}


def main():
    tables = {
        "test": read_biom(f"{REF}/src/test.biom"),
        "crawford": read_biom(f"{REF}/ci/crawford.biom"),
        "absquant": read_biom(f"{REF}/src/testdata/absquant.biom"),
    }
    trees = {}
    for key, p in [("test", "src/test.tre"), ("test_faith", "src/test_faith.tre"),
                   ("test_faith_shear", "src/test_faith_shear.tre"), ("crawford", "ci/crawford.tre"),
                   ("absquant", "src/testdata/absquant.tree")]:
        with open(f"{REF}/{p}") as fh:
            trees[key] = fh.read()
    faith = []
    with open(f"{REF}/ci/test.faith.exp") as fh:
        for ln in fh:
            if ln.strip():
                a, b = ln.split("\t")
                faith.append([a, b.strip()])
    expected = {
        "crawford_unweighted_fp64": read_dm(f"{REF}/ci/test.ssu.exp"),
        "crawford_unweighted_fp64_vaw": read_dm(f"{REF}/ci/test.ssu_vaw.exp"),
        "crawford_faith": faith,
        "absquant_wn_raw": read_dm_ws(f"{REF}/src/testdata/test.absquant.wn.dm"),
        "absquant_wn_normalized": read_dm_ws(f"{REF}/src/testdata/test.absquant_not.wn.dm"),
    }
    for src, dst in [("src/test.biom", "test.biom"), ("ci/crawford.biom", "crawford.biom"),
                     ("src/testdata/absquant.biom", "absquant.biom")]:
        shutil.copyfile(f"{REF}/{src}", os.path.join(OUT, dst))
    for name, obj in [("tables.json", tables), ("trees.json", trees), ("expected.json", expected)]:
        with open(os.path.join(OUT, name), "w") as fh:
            json.dump(obj, fh, indent=1)


if __name__ == "__main__":
    main()
