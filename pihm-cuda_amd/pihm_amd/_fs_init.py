"""Imported by the worker server (multiprocessing's forkserver) when it starts.

The server saves its SIGCHLD and SIGINT handlers and restores them in every worker it
forks.  A handler installed outside Python (by a preloaded profiler library, for instance)
reads back as None, which cannot be restored, and every worker then dies at start.  Such
handlers are replaced here, in the server process only, by Python's defaults."""
import signal

for _sig, _handler in ((signal.SIGCHLD, signal.SIG_DFL), (signal.SIGINT, signal.default_int_handler)):
    if signal.getsignal(_sig) is None:
        signal.signal(_sig, _handler)
