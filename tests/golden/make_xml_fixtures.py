"""Copy the example configuration files the reference ships (data: XML
parameter files of the live mode) into tests/golden, for the Conf / SimParam
/ OtfProvider / channel-registry tests.  Run from the repository root where
the reference tree is present; the copies are committed."""
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

for name in ("vigor-fastsim-example.xml", "vigor-omx-config-linux.xml"):
    shutil.copyfile(os.path.join(REF, name), os.path.join(HERE, name))
    print("copied", name)
