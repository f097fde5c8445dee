"""MI355X-native ApplicationClassification matcher (data graph x pattern graph -> per-vertex log-probabilities)."""
