"""Copies the reference test's output files used as fixtures here (data
only): tests/test-grains/spheres-settling-new-grains/Reference_results/
insert_position_x.dat, insert_coordinationNumber.dat and
insert_VelocityMaxMean.dat -- 20 spheres, 151 frames every 0.01 s -- and
the binary reload records insertA.result.bin with its _RFTable.txt and
_MatContact.xml.

Run in a checkout that has the reference tree at REFERENCE (default
/root/reference); the fixtures are committed, so the tests never read it.
"""
import os
import shutil
import sys

REFERENCE = os.environ.get("REFERENCE", "/root/reference")
SRC = os.path.join(REFERENCE, "tests", "test-grains", "spheres-settling-new-grains", "Reference_results")
FILES = ["insert_position_x.dat", "insert_coordinationNumber.dat", "insert_VelocityMaxMean.dat",
         "insertA.result.bin", "insert_RFTable.txt", "insert_MatContact.xml"]

if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    for f in FILES:
        shutil.copyfile(os.path.join(SRC, f), os.path.join(here, f))
    print("copied", len(FILES), "files", file=sys.stderr)
