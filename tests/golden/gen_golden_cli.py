#!/usr/bin/env python3
"""Writes tests/golden/qcgd_cli_*.txt: the output of the reference's QCGD
command-line flags, statistics and graph printing on a few runs.

oracle/_ref/ref_cli is oracle/ref_cli.cpp compiled against the reference
headers in place (oracle/Makefile); each fixture is its stdout for one
argument string (listed in qcgd_cli.json).  Run it where the reference
sources exist:

    make -C oracle ref && python tests/golden/gen_golden_cli.py

The runs stay below every truncation cap, so the states do not depend on
how ties at a cut are broken.
"""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "ref_cli")

RUNS = {
    # two iterations forward and back: the reversed run returns to the
    # initial three graphs; two initial specs (n_graphs, real, imag)
    "mixed": "2,seed=7,reversed_n_iter=2|3;2,n_graphs=2,real=0.6,imag=0.8|"
             "step;split_merge,theta=0.25,phi=0.1,xi=0.2;coin,theta=0.3;erase_create,theta=0.4,n_iter=2",
    # tolerance / align / truncation-mode / load-balancing flags, default angles
    "flags": "2,seed=11,tolerance=1e-12,align=4,simple_truncate=1,load_balancing_bucket_per_thread=8|4|"
             "erase_create,theta=0.2,phi=0.5;step;split_merge,xi=0.1;reversed_step",
    # print cut after 4 graphs
    "cut": "1,seed=5,max_print_num_graphs=4|3|split_merge;coin",
}


def main():
    for name, arg in RUNS.items():
        out = subprocess.run([REF_CLI, arg], capture_output=True, text=True, check=True, timeout=600).stdout
        with open(os.path.join(HERE, f"qcgd_cli_{name}.txt"), "w") as f:
            f.write(out)
    with open(os.path.join(HERE, "qcgd_cli.json"), "w") as f:
        json.dump(RUNS, f, indent=1)


if __name__ == "__main__":
    main()
