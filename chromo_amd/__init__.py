"""chromo_amd -- chromo's Monte Carlo inner loop on AMD Instinct MI355X (gfx950).

Drop-in host API (names and signatures of JosephWakim/chromo):
    chromo_amd.polymers  : PolymerBase, SSWLC, Chromatin, SSTWLC
    chromo_amd.binders   : ReaderProtein, hp1, prc1, null_reader, make_binder_collection
    chromo_amd.fields    : UniformDensityField, NullField
    chromo_amd.mc        : mc_sim, mc_step, polymer_in_field,
                           continue_polymer_in_field_simulation, get_amplitude_bounds, ...
    chromo_amd.util      : mu_schedules, poly_paths, poly_stat
Below it, libchromo_hip.so (include/chromo_hip.h) runs every proposal, energy
change, Metropolis test and density update in FP64 HIP kernels.  Extensions:
    chromo_amd.engine    : DeviceReplicas, many independent replicas in one persistent
                           launch, with checkpoint / restore
    chromo_amd.energy    : the dE-consistent energy, its adjoint and reweighting
    chromo_amd.autotune  : launch-shape autotuner
"""
__version__ = "0.1.0"
