"""MI355X-native meshless finite-volume (MFV) time step."""
