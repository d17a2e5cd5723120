"""MI355X-native two-site finite vMPS (DMRG) hot path of GraceQ/MPS2."""
