"""MI355X-native per-timestep force and integration path for the surfactant
droplet MD code (jhalverson/surfactant_enhanced_spreading)."""
