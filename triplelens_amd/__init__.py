"""MI355X-native extended-source light-curve engine (triple-lens IRS)."""
