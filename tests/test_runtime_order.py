"""The C-ABI library and torch share one HIP runtime whatever the import order (host only: library load maps)."""
import subprocess
import sys


def test_one_hip_runtime_when_library_loads_first():
    code = ("import ctypes\nfrom cerebrosim_amd import _native\n_native.load()\nimport torch\n"
            "maps = open('/proc/self/maps').read()\n"
            "paths = {l.split()[-1] for l in maps.splitlines() if 'libamdhip64' in l}\n"
            "print(len(paths))\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True,
                         cwd=__file__.rsplit("/tests/", 1)[0])
    assert out.stdout.strip().splitlines()[-1] == "1", out.stdout
