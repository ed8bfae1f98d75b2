"""MI355X-native Stokes hot path of M3TET_SPH (assembly -> Patera -> MGPCG)."""
