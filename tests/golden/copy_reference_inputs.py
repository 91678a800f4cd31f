"""Copies the data files of the reference's HIT test case that the tests use
as fixtures (run in a checkout that has the reference; the tests read only the
copies):

- samtaney_d4.in: the input deck tests/hit/Samtaney_D4/MPI_3DCompact.in
  (options-parser tests; the deck as shipped lacks the [LES] section);
- samtaney_d4_turbdata.txt: rows 1, 13, 25, 47, 78 of
  tests/hit/Samtaney_D4/turbdata.out (the published 128^3 statistics at
  t/tau = 0, 1, 2, 4, 7; columns t, tau, Re_lambda, M_t, eps, u', eta_K, rho',
  dil', omega', <rho u.u>).

    python tests/golden/copy_reference_inputs.py [/root/reference]
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    case = os.path.join(ref, "tests", "hit", "Samtaney_D4")
    with open(os.path.join(case, "MPI_3DCompact.in")) as f:
        deck = f.read()
    with open(os.path.join(HERE, "samtaney_d4.in"), "w") as f:
        f.write(deck)
    with open(os.path.join(case, "turbdata.out")) as f:
        rows = f.read().splitlines()
    keep = [1, 13, 25, 47, 78]
    with open(os.path.join(HERE, "samtaney_d4_turbdata.txt"), "w") as f:
        f.write("# line of tests/hit/Samtaney_D4/turbdata.out, then its 11 columns\n")
        for n in keep:
            f.write(f"{n} {rows[n - 1].strip()}\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
