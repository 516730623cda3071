"""MI355X (gfx950) backend for VOILA's data-parallel operator path."""
