"""MI355X-native exhaustive pairwise epistasis scan (the epiGPU pair-scan path).

Native code: libepiscan.so (HIP, gfx950) behind include/episcan.h and
libepihost.so (C host driver) behind include/epihost.h. This package holds
their sources, the in-tree build (epigpu_amd/lib/) and ctypes bindings.
"""
from .scan import (  # noqa: F401
    FORM_MFMA, FORM_PLAIN, HIT_DTYPE, Batch, EpiError, Scanner, f_isf, f_sf, layout, pack, perform_scan,
    read_container, read_plink, synth_genotypes, synth_phenotype, synth_plant, unpack, write_container,
    write_ped_map,
)
