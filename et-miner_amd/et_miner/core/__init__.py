"""Subpackage of et_miner."""
