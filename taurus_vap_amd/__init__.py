"""MI355X-native PNVAP projected-gradient path for TAURUS_vap."""
