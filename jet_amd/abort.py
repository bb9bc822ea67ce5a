"""Error type of the package, mirroring Jet::Exception (include/jet/Abort.hpp:51-97)."""


class JetError(RuntimeError):
    """Raised where the reference throws Jet::Exception.  The message contains
    the reference's own error text so that message checks written against the
    reference (`Contains("...")`) hold unchanged."""

    def __init__(self, message: str):
        super().__init__(f"Error in Jet: {message}")
        self.message = message
