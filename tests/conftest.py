import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vapoursynth-bm3dcuda_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    # the CPU oracle is a checker; build it if this tree has not been built yet
    lib = os.path.join(ROOT, "oracle", "_build", "libbm3d_oracle.so")
    if not os.path.exists(lib):
        subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "build_oracle.py")], check=True)
