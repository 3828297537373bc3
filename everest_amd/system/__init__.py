"""Command-line entry points (reference: system/single-query.py,
system/batch-query.py, system/filter.py, system/process-*.sh)."""
