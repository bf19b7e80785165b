"""Correlator classes with the constructor/correlate API of dynamix.correlator."""
