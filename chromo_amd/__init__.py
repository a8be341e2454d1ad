"""chromo_amd -- chromo's Monte Carlo inner loop on AMD Instinct MI355X (gfx950).

Drop-in host API (names and signatures of JosephWakim/chromo):
    chromo_amd.polymers  : PolymerBase, SSWLC, Chromatin
    chromo_amd.binders   : ReaderProtein, hp1, prc1, null_reader, make_binder_collection
    chromo_amd.fields    : UniformDensityField, NullField
    chromo_amd.mc        : mc_sim, mc_step, polymer_in_field, get_amplitude_bounds, all_moves, ...
Below it, libchromo_hip.so (include/chromo_hip.h) runs every proposal, energy
change, Metropolis test and density update in FP64 HIP kernels.  Extensions:
    chromo_amd.batch     : many independent replicas in one persistent launch
"""
__version__ = "0.1.0"
