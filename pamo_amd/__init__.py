"""MI355X-native PaMO stage-2 simplification and stage-3 safe projection."""
