"""CPU oracle for the petch etch-loop hot path — TEST INFRASTRUCTURE ONLY.

These modules restate the reference algorithm (sdelaurentiis123/plasma-etching-code,
src/petch/threed.py, belen.py, chemistry.py), each function citing the reference lines it
follows. They exist to check the HIP implementation: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import them, and nothing in the product package does.

    levelset.py   initial SDF (closed form), upwind advection, FSM Eikonal reinit, depth metrics
    chem.py       Belen coverages, surface_rate_belen / langmuir, yields, angular factors
    mesh.py       edge adjacency, flux smoothing, normalisation, mesh attributes, gas normals
    radiosity.py  form-factor assembly and the FP64 Jacobi / direct radiosity solves
    native.py     ctypes binding of the C restatement (petch_oracle.c): Philox, sources,
                  marching cubes, BVH-free and BVH ray queries, ion / neutral / legacy traces
    build.py      compiles petch_oracle.c with gcc (-fopenmp) into oracle/_build/
"""
