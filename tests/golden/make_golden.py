"""Write the golden fixtures in tests/golden/ from the reference's own code.

Run in a container that holds the reference sources (read-only):

    python tests/golden/make_golden.py

It compiles the reference's OpenMP backend and whole program from their
sources with oracle/Makefile (`make -C oracle ref`, outputs in oracle/_ref/),
then records inputs and outputs as small .npz fixtures:

  golden_precomp.npz  src/precomp.c:9-108 on a random model (2000 cells).
  golden_source.npz   src/calculate_source.c:10-17 for 3 dt values.
  golden_boundary.npz src/boundary.c:7-104 on the 2x2x2 split of a 26x24x28
                      global grid (each rank's local box with its ghost cells).
  golden_cell.npz     one src/openmp_propagate.c:dc_propagate call over the
                      compute box of random fields on a random TTI model.
  golden_run.npz      40 steps of the one-rank loop of src/worker.c:466-511
                      (source, the six boundary slabs + interior, swap) driven
                      through the reference's dc_propagate, dc_calculate_source
                      and box copies, on a random heterogeneous model.
  golden_dc_*.npz     the reference program bin/dc (src/main.c) under mpirun
                      with 1 and 8 ranks; the .dc output (pc then qc, global
                      x-fastest float32) plus the command line.
  golden_dims.npz     MPI_Dims_create(n, 3, {0,0,0}) for n = 1..4096, the
                      call src/main.c:89-90 makes, from the MPICH library
                      the reference links (its version string recorded).

`python tests/golden/make_golden.py --dims` writes golden_dims.npz alone.

Tests read these fixtures and the oracle.  Of the reference itself only one
compiled part leaves this container: oracle/_ref/libdcref.so, its OpenMP
dc_propagate built from its sources by oracle/Makefile, which travels with the
tree and is the CPU baseline bench.py times beside the GPU.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import oracle as O  # noqa: E402

MPIRUN = "/opt/conda/bin/mpirun"
MPI_HOME = "/opt/conda"

DIMS_PROGRAM = r"""
#include <mpi.h>
#include <stdio.h>
int main(int argc, char **argv) {
  char version[MPI_MAX_LIBRARY_VERSION_STRING];
  int len = 0;
  MPI_Init(&argc, &argv);
  MPI_Get_library_version(version, &len);
  printf("%s\n", version);
  for (int n = 1; n <= 4096; ++n) {
    int d[3] = {0, 0, 0};
    MPI_Dims_create(n, 3, d);
    printf("%d %d %d %d\n", n, d[0], d[1], d[2]);
  }
  MPI_Finalize();
  return 0;
}
"""


def write_dims():
    """golden_dims.npz: the MPI library's own MPI_Dims_create outputs."""
    with tempfile.TemporaryDirectory() as td:
        src, exe = os.path.join(td, "dims.c"), os.path.join(td, "dims")
        with open(src, "w") as f:
            f.write(DIMS_PROGRAM)
        lib = os.path.join(MPI_HOME, "lib")
        subprocess.run(["gcc", "-O1", "-I", os.path.join(MPI_HOME, "include"),
                        src, "-o", exe, "-L", lib, "-lmpi",
                        "-Wl,-rpath," + lib], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True,
                             timeout=120).stdout.splitlines()
    rows = [line.split() for line in out if line[:1].isdigit()]
    version = next(line for line in out if not line[:1].isdigit()).strip()
    a = np.array(rows, np.int32)
    np.savez_compressed(os.path.join(HERE, "golden_dims.npz"), n=a[:, 0],
                        dims=a[:, 1:], mpi_version=np.array(version))


def random_model(rng, shape):
    vpz = rng.uniform(1500.0, 4500.0, shape).astype(np.float32)
    vsv = (0.5 * vpz).astype(np.float32)
    eps = rng.uniform(0.0, 0.3, shape).astype(np.float32)
    dlt = (rng.uniform(0.0, 1.0, shape) * np.minimum(eps, 0.2)).astype(np.float32)
    theta = rng.uniform(0.0, np.pi / 3, shape).astype(np.float32)
    phi = rng.uniform(0.0, np.pi, shape).astype(np.float32)
    return dict(vpz=vpz, vsv=vsv, epsilon=eps, delta=dlt, phi=phi, theta=theta)


def model_coef(m):
    return O.ref_precomp(m["vpz"], m["vsv"], m["epsilon"], m["delta"], m["phi"],
                         m["theta"])


def slabs(sizes):
    """The region calls of src/worker.c:321-402 (boundaries, then interior)."""
    r = 4
    sx, sy, sz = sizes
    if not (sx >= 4 * r and sy >= 4 * r and sz >= 4 * r):
        return [((r, r, r), (sx - r, sy - r, sz - r))]
    out = []
    for side in (0, 1):
        out.append(((r if side == 0 else sx - 2 * r, r, r),
                    (2 * r if side == 0 else sx - r, sy - r, sz - r)))
    for side in (0, 1):
        out.append(((2 * r, r if side == 0 else sy - 2 * r, r),
                    (sx - 2 * r, 2 * r if side == 0 else sy - r, sz - r)))
    for side in (0, 1):
        out.append(((2 * r, 2 * r, r if side == 0 else sz - 2 * r),
                    (sx - 2 * r, sy - 2 * r, 2 * r if side == 0 else sz - r)))
    out.append(((2 * r,) * 3, (sx - 2 * r, sy - 2 * r, sz - 2 * r)))
    return out


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"],
                   check=True)
    rng = np.random.default_rng(20261203)

    # -- precompute -------------------------------------------------------
    m = random_model(rng, (1, 1, 2000))
    coef = model_coef(m)
    np.savez_compressed(os.path.join(HERE, "golden_precomp.npz"),
                        **{k: v.ravel() for k, v in m.items()},
                        **{n: c.ravel() for n, c in zip(O.COEF_NAMES, coef)})

    # -- Ricker source ----------------------------------------------------
    dts = np.array([1e-3, 2e-3, 2.5e-3], np.float32)
    tab = np.array([[O.ref_source(float(dt), i) for i in range(400)]
                    for dt in dts], np.float32)
    np.savez_compressed(os.path.join(HERE, "golden_source.npz"), dt=dts,
                        amp=tab)

    # -- random boundary on a 2x2x2 split --------------------------------
    prob = (10, 8, 12)
    absorb = 4
    gl = tuple(p + 2 * absorb + 8 for p in prob)
    part = tuple((g - 8) // 2 for g in gl)
    rec = {}
    for wz in range(2):
        for wy in range(2):
            for wx in range(2):
                w = (wx, wy, wz)
                loc = tuple(part[a] + (gl[a] - 8) % 2 * (w[a] == 1) + 8
                            for a in range(3))
                start = tuple(w[a] * part[a] for a in range(3))
                vpz = np.full(loc[::-1], 3000.0, np.float32)
                vsv = np.full(loc[::-1], 1296.1481, np.float32)
                seed = O.ref_random_boundary(loc, gl, start, prob, 4, absorb,
                                             vpz, vsv, 0)
                key = f"{wx}{wy}{wz}"
                rec[f"vpz_{key}"] = vpz
                rec[f"vsv_{key}"] = vsv
                rec[f"start_{key}"] = np.array(start)
                rec[f"seed_{key}"] = np.array(seed, np.uint64)
    np.savez_compressed(os.path.join(HERE, "golden_boundary.npz"),
                        problem=np.array(prob), absorb=np.array(absorb),
                        global_sizes=np.array(gl), **rec)

    # -- one dc_propagate call on random fields --------------------------
    shape = (12, 13, 14)  # (sz, sy, sx)
    m = random_model(rng, shape)
    coef = model_coef(m)
    f0 = [rng.standard_normal(shape).astype(np.float32) for _ in range(4)]
    f = [a.copy() for a in f0]
    sx, sy, sz = shape[2], shape[1], shape[0]
    O.ref_propagate((4, 4, 4), (sx - 4, sy - 4, sz - 4), f[0], f[1], f[2],
                    f[3], coef, 20.0, 17.5, 22.5, 1.5e-3)
    np.savez_compressed(os.path.join(HERE, "golden_cell.npz"),
                        dx=np.float32(20.0), dy=np.float32(17.5),
                        dz=np.float32(22.5), dt=np.float32(1.5e-3),
                        pp0=f0[0], pc0=f0[1], qp0=f0[2], qc0=f0[3],
                        pp1=f[0], qp1=f[2], **m)

    # -- 40-step one-rank loop through the reference's functions ---------
    shape = (18, 20, 22)
    m = random_model(rng, shape)
    coef = model_coef(m)
    sizes = (shape[2], shape[1], shape[0])
    f = [np.zeros(shape, np.float32) for _ in range(4)]
    src = (sizes[0] // 2, sizes[1] // 2, sizes[2] // 2)
    src_index = src[0] + src[1] * sizes[0] + src[2] * sizes[0] * sizes[1]
    dt, steps = np.float32(1.5e-3), 40
    for it in range(steps):
        s = np.float32(O.ref_source(float(dt), it))
        f[1].reshape(-1)[src_index] += s
        f[3].reshape(-1)[src_index] += s
        for a, b in slabs(sizes):
            O.ref_propagate(a, b, f[0], f[1], f[2], f[3], coef, 20.0, 20.0,
                            20.0, float(dt))
        f = [f[1], f[0], f[3], f[2]]
    np.savez_compressed(os.path.join(HERE, "golden_run.npz"), dt=dt,
                        steps=np.array(steps), src_index=np.array(src_index),
                        dx=np.float32(20.0), pc=f[1], qc=f[3], **m)

    # -- reference program, 1 and 8 ranks ---------------------------------
    cases = {
        "a": dict(size=(12, 14, 10), a=4, d=20.0, dt="0.002", tmax="0.08"),
        # 0.075f / 0.0025f rounds above 30, so ceil gives 31 iterations
        "b": dict(size=(10, 9, 11), a=3, d=25.0, dt="0.0025", tmax="0.075"),
    }
    with tempfile.TemporaryDirectory() as td:
        for name, c in cases.items():
            for np_ in (1, 8):
                out = os.path.join(td, f"{name}{np_}.dc")
                cmd = [MPIRUN, "-np", str(np_), O.REF_BIN_PATH,
                       "--size-x", str(c["size"][0]),
                       "--size-y", str(c["size"][1]),
                       "--size-z", str(c["size"][2]),
                       "--dx", str(c["d"]), "--dy", str(c["d"]),
                       "--dz", str(c["d"]), "--dt", c["dt"], "-t", c["tmax"],
                       "-a", str(c["a"]), "-o", out]
                env = dict(os.environ, OMP_NUM_THREADS="1")
                subprocess.run(cmd, check=True, env=env, capture_output=True,
                               timeout=300)
                gl = [s + 2 * c["a"] + 8 for s in c["size"]]
                data = np.fromfile(out, np.float32).reshape(2, gl[2], gl[1],
                                                            gl[0])
                np.savez_compressed(
                    os.path.join(HERE, f"golden_dc_{name}_np{np_}.npz"),
                    pc=data[0], qc=data[1], size=np.array(c["size"]),
                    absorption=np.array(c["a"]), d=np.float32(c["d"]),
                    dt=np.array(c["dt"]), tmax=np.array(c["tmax"]),
                    ranks=np.array(np_), argv=np.array(" ".join(cmd[4:])))
    write_dims()
    for fn in sorted(os.listdir(HERE)):
        if fn.endswith(".npz"):
            print(fn, os.path.getsize(os.path.join(HERE, fn)))


if __name__ == "__main__":
    if sys.argv[1:] == ["--dims"]:
        write_dims()
    else:
        main()
