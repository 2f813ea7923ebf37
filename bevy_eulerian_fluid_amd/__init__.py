"""MI355X-native per-step Eulerian fluid pipeline (bevy_eulerian_fluid's hot path)."""
