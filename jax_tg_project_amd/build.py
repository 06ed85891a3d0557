"""Build recipe for libtgdns.so (gfx950), in-tree next to this file.

    python -m jax_tg_project_amd.build

The kernels of each (precision, grid size) are one translation unit
(csrc/tg_inst.hip compiled with -DTG_INST_PREC / -DTG_INST_N); the
direct-sum path of every other size is csrc/tg_generic.hip; the C ABI is
csrc/tgdns.hip.  The objects compile in parallel and link into one shared
library; csrc/tg_cli.cpp links against it into the native command-line
solver taylor_green_hip.  Every call recompiles everything.
"""
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtgdns.so")
CLI = os.path.join(HERE, "taylor_green_hip")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # keep the compiler from emitting packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def grid_sizes():
    """The TG_GRID_SIZES list of csrc/tg_common.hpp."""
    with open(os.path.join(CSRC, "tg_common.hpp")) as f:
        m = re.search(r"#define TG_GRID_SIZES\(X\)((?:\s*X\(\d+\))+)", f.read())
    if not m:
        raise RuntimeError("TG_GRID_SIZES not found in tg_common.hpp")
    return [int(v) for v in re.findall(r"X\((\d+)\)", m.group(1))]


def _jobs():
    n = os.environ.get("MAX_JOBS")
    if n and n.isdigit():
        return max(1, int(n))
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    return cmd, res


def build(verbose=True, out=OUT):
    """Compile every translation unit and link libtgdns.so (always recompiles)."""
    cc = hipcc()
    tmpdir = tempfile.mkdtemp(prefix="tgdns_build_")
    try:
        units = [(os.path.join(CSRC, "tgdns.hip"), os.path.join(tmpdir, "tgdns.o"), []),
                 (os.path.join(CSRC, "tg_generic.hip"), os.path.join(tmpdir, "tg_generic.o"), []),
                 (os.path.join(CSRC, "tg_ens_mixed.hip"), os.path.join(tmpdir, "tg_ens_mixed.o"), [])]
        # largest grids first: they take longest
        for n in sorted(grid_sizes(), reverse=True):
            for prec in (64, 32):
                units.append((os.path.join(CSRC, "tg_inst.hip"), os.path.join(tmpdir, f"inst_{prec}_{n}.o"),
                              [f"-DTG_INST_PREC={prec}", f"-DTG_INST_N={n}"]))
        cmds = [[cc, *HIPCC_FLAGS, *defs, "-c", "-o", obj, src] for src, obj, defs in units]
        if verbose:
            print(f"compiling {len(cmds)} units with {min(_jobs(), len(cmds))} jobs: " + " ".join(cmds[0]), flush=True)
        failed = []
        with cf.ThreadPoolExecutor(max_workers=min(_jobs(), len(cmds))) as ex:
            for cmd, res in ex.map(_run, cmds):
                if res.returncode != 0:
                    failed.append((cmd, res))
        if failed:
            for cmd, res in failed:
                sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed for {len(failed)} unit(s)")
        tmp = out + ".tmp"
        link = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [obj for _, obj, _ in units]
        cmd, res = _run(link)
        if res.returncode != 0:
            sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
            raise RuntimeError(f"link failed with code {res.returncode}")
        os.replace(tmp, out)
        if out == OUT:
            # the native command-line solver, linked against the library next to it
            cli = [cc, "-O2", "-std=c++17", "-o", CLI + ".tmp", os.path.join(CSRC, "tg_cli.cpp"),
                   "-L" + HERE, "-ltgdns", "-Wl,-rpath,$ORIGIN"]
            cmd, res = _run(cli)
            if res.returncode != 0:
                sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
                raise RuntimeError(f"building taylor_green_hip failed with code {res.returncode}")
            os.replace(CLI + ".tmp", CLI)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return out


if __name__ == "__main__":
    build()
