#!/usr/bin/env python3
"""Copies the input data of the reference's 1R69 single-memory example
(examples/examples_new/1R69_single_memory) into tests/golden/deck_1r69/, so
that the deck front end can run the shipped system where /root/reference is
absent.  Data files only: the atom_style awsemmd data file, the coeff file
(fragment memory in table mode on the decks' 0-50 A by 0.01 A grid), the
sequence, ssweight, the single memory (the native structure's .gro) with its
.mem list and gamma file, and the parameter tables the coeff names.  The
input script is not copied: the tests write their own, with the example's
settings (timestep 5 fs, Langevin 500 -> 300 K, damping 1000 fs).

    python tests/golden/make_deck_1r69.py
"""
import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/examples/examples_new/1R69_single_memory"
DST = os.path.join(HERE, "deck_1r69")
FILES = ["data.1r69", "fix_backbone_coeff.data", "extended.seq", "ssweight", "fragsLAMW.mem.single", "1r69a.gro",
         "seq.gamma", "anti_HB", "anti_NHB", "anti_one", "para_HB", "para_one", "gamma.dat", "burial_gamma.dat"]


def main():
    os.makedirs(DST, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(SRC, f), os.path.join(DST, f))
    print("copied", len(FILES), "files to", DST)


if __name__ == "__main__":
    main()
