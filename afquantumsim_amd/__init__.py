"""MI355X-native state-vector engine with the aqs:: API of afQuantumSim."""
