"""MI355X (gfx950) path-tracing core with lajolla's render API."""
