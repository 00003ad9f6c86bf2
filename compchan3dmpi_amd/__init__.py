"""MI355X-native time step of the compressible channel DNS chanCompr."""
