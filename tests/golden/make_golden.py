"""Writes tests/golden/omp3_kernels.npz from the reference omp3 kernels.

The reference kernels are compiled from their own sources by `make -C oracle
ref` into oracle/_ref/libomp3_{2d,3d}.so (serial, -ffp-contract=off).  For
each case in tests.kernel_cases.golden_cases() this stores the seeded inputs
and the reference's outputs, so the oracle can be checked against the
reference on machines where the reference is not present.

Run from the repository root:  python -m tests.golden.make_golden
"""
import numpy as np

from tests import kernel_cases as kc
from oracle.pyoracle import load_ref


def main():
    out = {}
    for case in kc.golden_cases():
        ref = load_ref(case.dims)
        if ref is None:
            raise SystemExit("build the reference first: make -C oracle ref")
        cid = kc.case_id(case)
        for k, v in case.fields.items():
            out[f"{cid}:in:{k}"] = np.asarray(v)
        res = kc.run_host(ref, "", case, ref=True)
        for k, v in res.items():
            out[f"{cid}:{k}"] = np.asarray(v)
    np.savez_compressed(kc.GOLDEN_FILE, **out)
    print(f"wrote {len(out)} arrays to {kc.GOLDEN_FILE}")


if __name__ == "__main__":
    main()
