"""CPU oracle for the LEX dynamics hot path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import it; the
product (lex_amd) never does.
"""
