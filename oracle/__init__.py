"""Oracle (test infrastructure only): CPU restatement of the reference hot path."""
