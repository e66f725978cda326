"""camp_amd: MI355X-native batched CAMP chemistry solver (package stub)."""
