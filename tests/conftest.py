import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "yatscf-df_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
XYZ = os.path.join(GOLDEN, "xyz")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def system(name):
    import basis
    return basis.System(os.path.join(PKG, "data", "cc-pvdz", "cc-pvdz.gbs"),
                        os.path.join(PKG, "data", "cc-pVDZ-RI", "cc-pVDZ-RI.gbs"),
                        os.path.join(XYZ, name + ".xyz"))


def xyz_path(name):
    return os.path.join(XYZ, name + ".xyz")
