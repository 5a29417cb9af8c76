"""Equation backends selectable by the YAML ``equation:`` key."""
