"""Drop-in entry points with the reference's signatures (seqm/seqm_functions/*)."""
