"""Loading the native library before anything imports torch must still leave
one HIP runtime in the process (torch's): the loader imports torch first."""

import subprocess
import sys


def test_library_loads_torch_runtime_first():
    code = ("import sys\n"
            "from up3d_amd import _native\n"
            "assert 'torch' not in sys.modules\n"
            "_native.lib()\n"
            "assert 'torch' in sys.modules\n"
            "maps = open('/proc/self/maps').read()\n"
            "paths = {l.split()[-1] for l in maps.splitlines() if 'libamdhip64' in l}\n"
            "assert len(paths) == 1, paths\n"
            "print(next(iter(paths)))\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=300, cwd=__file__.rsplit("/tests/", 1)[0])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "torch" in r.stdout
