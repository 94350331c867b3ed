"""MI355X-native numeric engine for NASOQ's LBL sparse symmetric-indefinite solver."""
