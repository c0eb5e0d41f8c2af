"""MI355X-native PN-CSD forward/adjoint solver for EPMA k-ratios."""
