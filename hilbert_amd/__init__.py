"""MI355X-native fixed-orbital v2RDM (DQG) boundary-point SDP solver."""
