"""The product path has no CPU fallback: the package never imports the
oracle, and without the HIP library every engine call fails loudly."""
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_package_does_not_import_the_oracle():
    pat = re.compile(r"^\s*(import\s+oracle|from\s+oracle\b)", re.M)
    for dirpath, _, files in os.walk(os.path.join(ROOT, "fsig_amd")):
        for name in files:
            if name.endswith(".py"):
                text = open(os.path.join(dirpath, name)).read()
                assert not pat.search(text), os.path.join(dirpath, name)


def test_missing_hip_library_raises():
    code = (
        "import fsig_amd as F\n"
        "F.HIP_LIB_PATH = '/nonexistent/libfsig_hip.so'\n"
        "F._hip = None\n"
        "try:\n"
        "    F.hip_lib()\n"
        "except FileNotFoundError as e:\n"
        "    print('raised:', e)\n"
        "else:\n"
        "    print('no error')\n"
    )
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    assert "raised:" in r.stdout and "no CPU fallback" in r.stdout
