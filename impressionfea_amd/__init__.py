"""MI355X-native 3D rigid-viscoplastic solve step (ImpressionFEA's Forging3DSolver hot path)."""
