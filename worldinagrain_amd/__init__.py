"""worldinagrain's liquid step (FLIP/APIC + MG-PCG) for AMD MI355X.

- `worldinagrain_amd.gpu`: the HIP world (libwig.so, include/wig.h);
- `worldinagrain_amd.host`: the host side of js/sim.js's liquid step;
- `worldinagrain_amd.scenes`: initial worlds in the reference's cell format.
"""
