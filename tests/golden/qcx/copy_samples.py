"""Copies the reference's sample quantum-assembler programs (bra/qcx/*.qcx,
inputs of bra) next to this script as test fixtures.  Run in a checkout
that has the reference at /root/reference; the copies are committed."""
import os
import shutil

SRC = "/root/reference/bra/qcx"
HERE = os.path.dirname(os.path.abspath(__file__))

for name in ("hadamards08.qcx", "toffoli12.qcx", "adder6x2.qcx", "shor12.qcx"):
    shutil.copyfile(os.path.join(SRC, name), os.path.join(HERE, name))
    print("copied", name)
