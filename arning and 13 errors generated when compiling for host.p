// HIP kernels of the DQMC engine for gfx950 (MI355X).
